// k_turbulence.hip - vertical turbulence closure (turbulence_closure_v1.f90:9-492).
//
// Scheme 1 (Richardson step) and 2 (Pacanowski-Philander) are per-level algebra. Scheme 3
// solves the k (q2) and psi equations of the generic length scale closure as two
// tridiagonal systems per face column, one thread per column: the stability functions of
// the current state and the layer quantities as the elimination reaches each row, the two
// Thomas recurrences of tridiagonal_solver.f90, the length scale, the conversion to whole
// levels and the new diffusivities. The stability
// functions are the quasi-equilibrium ones of Galperin et al. (1988) or Kantha & Clayson
// (1994) (the reference's algebraic_stress is undefined). The face diffusivity is then
// averaged onto cells for the transport (Kv). Each value is computed by the same expression
// sequence as the CPU restatement.
#include "gomgpu_dev.h"

#define TK_A1 0.92
#define TK_A2 0.74
#define TK_B1 16.6
#define TK_B2 10.1
#define TK_C1 0.08
#define TK_C2 0.7
#define TK_C3 0.2
#define TLEV (GOM_LMAX + 2)

/* x**e for the GLS exponents: the ones the models use (0, +-1, 2, 3, +-0.5, 1.5) by exact
 * arithmetic, so host and device agree bit for bit; others through pow. */
__device__ __forceinline__ double gls_pow(double x, double e) {
  if (e == 0.0) return 1.0;
  if (e == 1.0) return x;
  if (e == -1.0) return 1.0 / x;
  if (e == 2.0) return x * x;
  if (e == 3.0) return x * x * x;
  if (e == 0.5) return sqrt(x);
  if (e == -0.5) return 1.0 / sqrt(x);
  if (e == 1.5) return x * sqrt(x);
  if (e == -1.5) return 1.0 / (x * sqrt(x));
  if (e == -3.0) return 1.0 / (x * x * x);
  return pow(x, e);
}

__device__ __forceinline__ void stability(int kind, double gh, double &sm, double &sh) {
  if (gh > 0.0233) gh = 0.0233;
  if (gh < -0.28) gh = -0.28;
  double b2 = kind == GOMGPU_STAB_KC ? TK_B2 * (1.0 - TK_C3) : TK_B2;
  double a2 = kind == GOMGPU_STAB_KC ? TK_A2 * (1.0 - TK_C2) : TK_A2;
  double h = TK_A2 * (1.0 - 6.0 * TK_A1 / TK_B1) / (1.0 - 3.0 * TK_A2 * gh * (6.0 * TK_A1 + b2));
  double m = (TK_A1 * (1.0 - 3.0 * TK_C1 - 6.0 * TK_A1 / TK_B1) + 9.0 * TK_A1 * (2.0 * TK_A1 + a2) * h * gh) /
             (1.0 - 9.0 * TK_A1 * TK_A2 * gh);
  sm = m;
  sh = h;
}

// algebraic_stress at one whole level: vd, td, qd, qd2 clipped to [bgdiff, diffmax]
__device__ __forceinline__ void alg_stress(const TurbParams &p, double q2, double xl, double rz, double &vd,
                                           double &td, double &qd, double &qd2) {
  double gh = rz * (xl * xl) / (2.0 * q2);
  double sm, sh;
  stability(p.stab, gh, sm, sh);
  double q = sqrt(2.0 * q2);
  double v = sm * q * xl, t = sh * q * xl, a, b;
  if (p.model == GOMGPU_GLS_MY || p.model == GOMGPU_GLS_KL) {
    a = 0.2 * q * xl;
    b = a;
  } else {
    a = v / p.schk;
    b = v / p.schpsi;
  }
  vd = dminf(p.diffmax, dmaxf(p.bgdiff, v));
  td = dminf(p.diffmax, dmaxf(p.bgdiff, t));
  qd = dminf(p.diffmax, dmaxf(p.bgdiff, a));
  qd2 = dminf(p.diffmax, dmaxf(p.bgdiff, b));
}

// Schemes 1 and 2 (turbulence_closure_v1.f90:75-141): per-level algebra, one thread per
// (level, face).
__global__ void k_turb_ri(DevMesh m, DevState s, DevWork w, DevDims d, TurbParams p) {
  long t = (long)gom_block() * blockDim.x + threadIdx.x;
  const int J = d.J, L = d.L;
  if (t >= (long)L * J) return;
  int j = (int)(t % J), k = (int)(t / J) + 1;
  int kt = s.top_layer_at_face[j];
  if (kt == 0) return;
  int kb = m.bottom_layer_at_face[j];
  if (k < kb || k > kt) return;
  const double g = p.g;
  double vmax = 0.0, vmin = 0.0;
  if (p.flag == 2) {
    double h = m.h_face[j];
    if (h <= p.h1_pp) {
      vmax = p.vdmax_pp1;
      vmin = p.vdmin_pp1;
    } else if (h < p.h2_pp) {
      vmax = p.vdmax_pp1 + (p.vdmax_pp2 - p.vdmax_pp1) * (h - p.h1_pp) / (p.h2_pp - p.h1_pp);
      vmin = p.vdmin_pp1 + (p.vdmin_pp2 - p.vdmin_pp1) * (h - p.h1_pp) / (p.h2_pp - p.h1_pp);
    } else {
      vmax = p.vdmax_pp2;
      vmin = p.vdmin_pp2;
    }
  }
  double ri;
  if (k == kt) {
    ri = 0.0;
  } else {
    double dzh = A0(s.dzhalf_face_new, k, j, J);
    double drhodz = (AL(s.rho_face, k + 1, j, J) - AL(s.rho_face, k, j, J)) / dzh;
    double bvf = -g * (drhodz / p.rho_o + g / 2250000.0);
    double dundz = (AL(s.un_face_new, k + 1, j, J) - AL(s.un_face_new, k, j, J)) / dzh;
    double dutdz = (AL(s.vn_face_new, k + 1, j, J) - AL(s.vn_face_new, k, j, J)) / dzh;
    double shear2 = dundz * dundz + dutdz * dutdz;
    shear2 = dmaxf(shear2, 1.0e-10);
    ri = dmaxf(bvf / shear2, 0.0);
  }
  double av, kv;
  if (p.flag == 1) {
    if (ri < 0.25) {
      av = 1.0e-3;
      kv = 1.0 * 1.0e-3;
    } else {
      av = 1.0e-5;
      kv = 1.0e-5;
    }
  } else {
    double tt = 1.0 + 5.0 * ri;
    av = vmax / (tt * tt) + vmin;
    kv = av / (1.0 + 5.0 * ri) + p.tdmin_pp;
  }
  A0(s.Av, k, j, J) = av;
  A0(w.turb_kvf, k, j, J) = kv;
}

// Scheme 3 (turbulence_closure_v1.f90:56-66, 158-427), one thread per face column:
// * the two stop checks of the layer quantities (:158-215);
// * the k and psi tridiagonal systems (Thomas in place in w.tk_q2tmp / w.tk_x, factors in
//   w.tri_gam and w.tk_shear), each row's inputs -- the level diffusivities of the current
//   k, l, the half-level k, l, the squared shear and (g/rho_o) drho/dz -- formed as the
//   top-down elimination reaches it; each row reads the level below it once and carries
//   the level above (its state, diffusivities, density and node velocities) from the row
//   before, where level kb-1 takes level kb's state;
// * k_turb_levels then completes the column.
__global__ void __launch_bounds__(256) GOM_WAVES(GOM_WAVES_TURB) k_turb_column(DevMesh m, DevState s, DevForcing f,
                                                                                DevWork w, DevDims d, TurbParams p,
                                                                                int *err) {
  int j = gom_block() * blockDim.x + threadIdx.x;
  const int J = d.J;
  if (j >= J) return;
  int kt = s.top_layer_at_face[j];
  if (kt == 0) return;
  int kb = m.bottom_layer_at_face[j];
  const double dt = p.dt;
  int a1 = TB(m.adj_cellnum_at_face, 1, j, J), a2 = TB(m.adj_cellnum_at_face, 2, j, J);
  double etam = a2 != 0 ? dmaxf(s.eta_cell_new[a1 - 1], s.eta_cell_new[a2 - 1]) : s.eta_cell_new[a1 - 1];
  double hf = m.h_face[j];
  int nq = kt - kb + 1;
  double cm3 = p.cmiu0 * p.cmiu0 * p.cmiu0;
  const int P = d.P;
  const int n1 = TB(m.nodenum_at_face, 1, j, J) - 1, n2 = TB(m.nodenum_at_face, 2, j, J) - 1;
  // the stop checks of every layer first, as the reference's pass over the layers precedes
  // its solve
  double xl_lo = A0(s.xl, kb - 1, j, J);
  for (int kl = kb; kl <= kt; kl++) {
    const double xl_hi = A0(s.xl, kl, j, J);
    double xlbt = (xl_lo + xl_hi) / 2.0;
    xl_lo = xl_hi;
    if (xlbt < p.xlmin2) latch(err, GOMGPU_ERR_TURB_XL, j + 1);
    double dzk = AL(s.dz_face_new, kl, j, J);
    double zctr = (kl == kt) ? p.MSL + etam - dzk / 2.0 : m.z_level[kl] - dzk / 2.0;
    double dists = p.MSL + etam - zctr, distb = zctr - (p.MSL - hf);
    double xlmax = (double)0.4f * dists * distb / (hf + etam);
    if (xlmax <= 0.0) latch(err, GOMGPU_ERR_TURB_DIST, j + 1);
  }
  // level kt's state, diffusivities and node velocities, carried down the rows: row k reads
  // level kl-1 once and takes level kl (and the density at kl+1) from the row above
  double dzk = AL(s.dz_face_new, kt, j, J), dzhk = A0(s.dzhalf_face_new, kt, j, J);
  double q2k = A0(s.q2, kt, j, J), xlk = A0(s.xl, kt, j, J);
  double rhok = AL(s.rho_face, kt, j, J), rhop = 0.0;  // density at kl and kl+1
  double avk, kvk, qdk, qd2k;
  alg_stress(p, q2k, xlk, 0.0, avk, kvk, qdk, qd2k);
  double uup = (A0(s.u_node, kt, n1, P) + A0(s.u_node, kt, n2, P)) / 2.0;
  double vup = (A0(s.v_node, kt, n1, P) + A0(s.v_node, kt, n2, P)) / 2.0;
  // ---- the k and psi equations (:233-276 and :281-347) in one top-down pass: the two
  // systems share their inputs and are independent, so each row's inputs are read once and
  // both eliminations advance together (rows top down; row k <-> layer kl = kt - k + 1).
  // The k system's factors go to w.tri_gam (by row), the psi system's to w.tk_shear (by
  // layer); solutions in w.tk_q2tmp and w.tk_x.
  const double cpow = gls_pow(p.cmiu0, p.rpub);
  double den = 0.0, cprev = 0.0, x = 0.0;     // k system
  double denp = 0.0, cprevp = 0.0, xp = 0.0;  // psi system
  for (int k = 1; k <= nq; k++) {
    const int kl = kt - k + 1, km = kl - 1;
    const double dzhm = A0(s.dzhalf_face_new, km, j, J);
    const double q2m = A0(s.q2, km, j, J), xlm = A0(s.xl, km, j, J);
    double dzm = 0.0, rhom = 0.0;
    double avm, kvm, qdm, qd2m;  // level kl-1; level kb-1 takes level kb's
    if (km >= kb) {
      dzm = AL(s.dz_face_new, km, j, J);
      rhom = AL(s.rho_face, km, j, J);
      alg_stress(p, q2m, xlm, p.g / p.rho_o * ((rhok - rhom) / dzhm), avm, kvm, qdm, qd2m);
    } else {
      avm = avk;
      kvm = kvk;
      qdm = qdk;
      qd2m = qd2k;
    }
    // layer kl (turbulence_closure_v1.f90:158-205)
    double q2bt = (q2m + q2k) / 2.0;
    double xlbt = (xlm + xlk) / 2.0;
    double udown = (A0(s.u_node, kl - 1, n1, P) + A0(s.u_node, kl - 1, n2, P)) / 2.0;
    double vdown = (A0(s.v_node, kl - 1, n1, P) + A0(s.v_node, kl - 1, n2, P)) / 2.0;
    double dudz = (uup - udown) / dzk, dvdz = (vup - vdown) / dzk;
    double shear = dudz * dudz + dvdz * dvdz;
    double drhodz;
    if (kl == kt || kl == kb) {
      drhodz = 0.0;
    } else {
      double rho_up = rhok + dzk / 2.0 / dzhk * (rhop - rhok);
      double rho_down = rhom + dzm / 2.0 / dzhm * (rhok - rhom);
      drhodz = (rho_up - rho_down) / dzk;
    }
    double rz = p.g / p.rho_o * drhodz;
    AL(w.tk_rz, kl, j, J) = rz;
    // k row
    double eleft = 0.0, eright = 0.0, a = 0.0, c = 0.0;
    if (k > 1) {
      eleft = -(qdk * dt / dzhk);
      a = eleft;
    }
    if (k < nq) {
      eright = -(qdm * dt / dzhm);
      c = eright;
    }
    double prod = (avk + avm) / 2.0 * shear;
    double buoy = (kvk + kvm) / 2.0 * rz;
    double diss = cm3 * sqrt(q2bt) / xlbt;
    double pplus, pminus;
    if (prod + buoy > 0.0) {
      pplus = prod + buoy;
      pminus = 0.0;
    } else {
      pplus = 0.0;
      pminus = -prod - buoy;
    }
    double b = dzk - eleft - eright + (diss + pminus / q2bt) * dt * dzk;
    double r = dzk * (q2bt + dt * pplus);
    if (k == 1) {
      if (b == 0.0) {
        latch(err, GOMGPU_ERR_TRIDIAG_PIVOT, j + 1);
        return;
      }
      den = b;
      x = r / den;
    } else {
      double gam = cprev / den;
      den = b - a * gam;
      if (den == 0.0) {
        latch(err, GOMGPU_ERR_TRIDIAG_PIVOT, j + 1);
        return;
      }
      x = (r - a * x) / den;
      AL(w.tri_gam, k, j, J) = gam;
    }
    AL(w.tk_q2tmp, kl, j, J) = x;
    cprev = c;
    // psi row
    double eleftp = 0.0, erightp = 0.0, ap = 0.0, cp = 0.0;
    if (k > 1) {
      eleftp = -(qd2k * dt / dzhk);
      ap = eleftp;
    }
    if (k < nq) {
      erightp = -(qd2m * dt / dzhm);
      cp = erightp;
    }
    double bb = dzk - eleftp - erightp;
    double cpsi2p;
    if (p.model == GOMGPU_GLS_MY || p.model == GOMGPU_GLS_KL) {
      double zctr = m.z_level[kl] - dzk / 2.0;
      double dists = dmaxf(p.MSL + etam - zctr, p.xlmin2);
      double distb = dmaxf(zctr - (p.MSL - hf), p.xlmin2);
      double r1 = xlbt / (double)0.4f / distb, r2 = xlbt / (double)0.4f / dists;
      double psi = 1.0 + (double)1.33f * (r1 * r1) + (double)0.25f * (r2 * r2);
      cpsi2p = psi * p.cpsi2;
    } else {
      cpsi2p = p.cpsi2;
    }
    double dissp = cpsi2p * cm3 * sqrt(q2bt) / xlbt;
    bb = bb + dt * dzk * dissp;
    double prodp = p.cpsi1 * (avk + avm) / 2.0 * shear;
    double cpsi3;
    if (p.model == GOMGPU_GLS_MY) {
      cpsi3 = (double)0.9f;
    } else if (rz > 0.0) {
      cpsi3 = 1.0;
    } else {
      switch (p.model) {
        case GOMGPU_GLS_KL: cpsi3 = (double)2.53f; break;
        case GOMGPU_GLS_KE: cpsi3 = (double)(-0.52f); break;
        case GOMGPU_GLS_KW: cpsi3 = (double)(-0.58f); break;
        default: cpsi3 = (double)0.1f; break;
      }
    }
    double buoyp = cpsi3 * (kvk + kvm) / 2.0 * rz;
    double pplusp, pminusp;
    if (prodp + buoyp > 0.0) {
      pplusp = prodp + buoyp;
      pminusp = 0.0;
    } else {
      pplusp = 0.0;
      pminusp = -prodp - buoyp;
    }
    bb = bb + pminusp / q2bt * dt * dzk;
    if (kl == kb || kl == kt) bb = bb + dt * (double)0.4f * p.rnub * qd2k / xlbt;
    double tmp = cpow * gls_pow(q2bt, p.rmub) * gls_pow(xlbt, p.rnub);
    double rp = dzk * tmp * (1.0 + dt * pplusp / q2bt);
    if (k == 1) {
      if (bb == 0.0) {
        latch(err, GOMGPU_ERR_TRIDIAG_PIVOT, j + 1);
        return;
      }
      denp = bb;
      xp = rp / denp;
    } else {
      double gam = cprevp / denp;
      denp = bb - ap * gam;
      if (denp == 0.0) {
        latch(err, GOMGPU_ERR_TRIDIAG_PIVOT, j + 1);
        return;
      }
      xp = (rp - ap * xp) / denp;
      AL(w.tk_shear, kl, j, J) = gam;
    }
    AL(w.tk_x, kl, j, J) = xp;
    cprevp = cp;
    avk = avm;
    kvk = kvm;
    qdk = qdm;
    qd2k = qd2m;
    uup = udown;
    vup = vdown;
    dzk = dzm;
    dzhk = dzhm;
    q2k = q2m;
    xlk = xlm;
    rhop = rhok;
    rhok = rhom;
  }
}

// The rest of scheme 3 (turbulence_closure_v1.f90:268-278, 351-427), one thread per face
// column, after k_turb_column's elimination: the back substitutions of both systems from
// the bottom row's solutions (w.tk_q2tmp, w.tk_x at layer kb), the length scale, k and l at
// whole levels and the new diffusivities. A kernel of its own so that neither half holds
// the other's registers; the surface and bottom values of k and l are formed again from
// the same inputs. A column whose elimination stopped at a zero pivot has failed the step
// (the error is latched) and its values here are not used.
__global__ void __launch_bounds__(256) GOM_WAVES(GOM_WAVES_TURB) k_turb_levels(DevMesh m, DevState s, DevForcing f,
                                                                                DevWork w, DevDims d, TurbParams p,
                                                                                int *err) {
  int j = gom_block() * blockDim.x + threadIdx.x;
  const int J = d.J;
  if (j >= J) return;
  int kt = s.top_layer_at_face[j];
  if (kt == 0) return;
  int kb = m.bottom_layer_at_face[j];
  double wn = f.wind_stress_normal[j], wt = f.wind_stress_tangnt[j];
  double q2fs = p.fs_coef * sqrt(wn * wn + wt * wt) / 2.0;
  q2fs = dmaxf(q2fs, p.q2min);
  double ubn = AL(s.un_face_new, kb, j, J), ubt = AL(s.vn_face_new, kb, j, J);
  double q2bot = p.fs_coef * s.Cdb[j] * sqrt(ubn * ubn + ubt * ubt) / 2.0;
  q2bot = dmaxf(q2bot, p.q2min);
  double xlfs = dmaxf(p.xlmin1, AL(s.dz_face_new, kt, j, J) / 2.0 * (double)0.4f);
  double xlbot = dmaxf(p.xlmin2, AL(s.dz_face_new, kb, j, J) / 2.0 * (double)0.4f);
  const int nq = kt - kb + 1;
  // ---- back substitutions, bottom up, both systems (:268-278, :351-368): the k solution
  // with its boundary values, then the length scale from the psi solution of the same row
  double x = AL(w.tk_q2tmp, kb, j, J), xp = AL(w.tk_x, kb, j, J);
  const double cpow_inv = gls_pow(p.cmiu0, -p.rpub);
  for (int k = nq; k >= 1; k--) {
    int kl = kt - k + 1;
    if (k < nq) x = AL(w.tk_q2tmp, kl, j, J) - AL(w.tri_gam, k + 1, j, J) * x;
    const double q2tmp = kl == kt ? q2fs : (kl == kb ? q2bot : dmaxf(x, p.q2min));
    AL(w.tk_q2tmp, kl, j, J) = q2tmp;
    if (k < nq) xp = AL(w.tk_x, kl, j, J) - AL(w.tk_shear, kl - 1, j, J) * xp;
    double q2l = dmaxf(xp, p.psimin);
    double rz = AL(w.tk_rz, kl, j, J);
    double xv;
    if (kl == kt)
      xv = xlfs;
    else if (kl == kb)
      xv = xlbot;
    else
      xv = gls_pow(q2l * cpow_inv * gls_pow(q2tmp, -p.rmub), 1.0 / p.rnub);
    if (rz < 0.0) {
      double upper = sqrt(-(double)0.56f * q2tmp / rz);
      xv = dminf(xv, upper);
    }
    double cq = p.cmiu0 * sqrt(q2tmp);
    double xl_max = cq * cq * cq / p.eps_min;
    // the solution of row k-1 still needs x, so the clipped length goes to the xl array
    A0(s.xl, kl, j, J) = dmaxf(p.xlmin2, dminf(xl_max, xv));
  }
  // whole levels (:372-390) and the new diffusivities (:397-427), descending from the
  // surface. Level k interpolates layers k and k+1: layer k is read before level k replaces
  // it, and layer k+1 (read before level k+1 replaced it) and the density at k+1 carry from
  // the level above. Each level's stability functions follow at once; the levels above kt
  // take level kt's k, l and diffusivities, those below kb level kb-1's k, l and level kb's
  // diffusivities.
  double vd, td, qd, qd2;
  alg_stress(p, q2fs, xlfs, 0.0, vd, td, qd, qd2);
  for (int k = kt + 1; k <= d.L; k++) {
    A0(s.q2, k, j, J) = q2fs;
    A0(s.xl, k, j, J) = xlfs;
    A0(s.Av, k, j, J) = vd;
    A0(w.turb_kvf, k, j, J) = td;
  }
  double q2b = AL(w.tk_q2tmp, kt, j, J), xb = A0(s.xl, kt, j, J), rho_b = AL(s.rho_face, kt, j, J);
  A0(s.q2, kt, j, J) = q2fs;
  A0(s.xl, kt, j, J) = xlfs;
  A0(s.Av, kt, j, J) = vd;
  A0(w.turb_kvf, kt, j, J) = td;
  for (int k = kt - 1; k >= kb; k--) {
    const double dzh = A0(s.dzhalf_face_new, k, j, J);
    const double r = AL(s.dz_face_new, k, j, J) / 2.0 / dzh;
    const double q2a = AL(w.tk_q2tmp, k, j, J), xa = A0(s.xl, k, j, J), rho_a = AL(s.rho_face, k, j, J);
    const double q = q2a + r * (q2b - q2a);
    const double xv = xa + r * (xb - xa);
    if (q < 0.0 || xv < p.xlmin2) {
      latch(err, GOMGPU_ERR_TURB_NEGATIVE, j + 1);
      return;
    }
    A0(s.q2, k, j, J) = q;
    A0(s.xl, k, j, J) = xv;
    alg_stress(p, q, xv, p.g / p.rho_o * ((rho_b - rho_a) / dzh), vd, td, qd, qd2);
    A0(s.Av, k, j, J) = vd;
    A0(w.turb_kvf, k, j, J) = td;
    q2b = q2a;
    xb = xa;
    rho_b = rho_a;
  }
  for (int k = 0; k < kb; k++) {
    A0(s.q2, k, j, J) = q2bot;
    A0(s.xl, k, j, J) = xlbot;
    if (k > 0) {
      A0(s.Av, k, j, J) = vd;
      A0(w.turb_kvf, k, j, J) = td;
    }
  }
}

// Face diffusivity -> cell Kv: mean over the cell's wet faces at the level clamped to each
// face's active range (1.0e-2 for a cell with no wet face); one thread per (level, cell).
__global__ void k_turb_kv(DevMesh m, DevState s, DevWork w, DevDims d) {
  // GOM_KV_LEVELS consecutive levels of one cell per thread: the cell's faces and their wet
  // ranges are read once for all of them
  const int E = d.E, J = d.J, L1 = d.L + 1;
  const int nch = (L1 + GOM_KV_LEVELS - 1) / GOM_KV_LEVELS;
  long t = (long)gom_block() * blockDim.x + threadIdx.x;
  if (t >= (long)nch * E) return;
  const int i = (int)(t % E), k0 = (int)(t / E) * GOM_KV_LEVELS;
  if (s.top_layer_at_element[i] == 0) return;
  const int N = m.tri_or_quad[i];
  int fc[4], fkt[4], fkb[4];
#pragma unroll
  for (int l = 0; l < 4; l++) {
    fc[l] = l < N ? TB(m.facenum_at_cell, l + 1, i, E) - 1 : 0;
    fkt[l] = l < N ? s.top_layer_at_face[fc[l]] : 0;
    fkb[l] = l < N ? m.bottom_layer_at_face[fc[l]] : 1;
  }
#pragma unroll
  for (int z = 0; z < GOM_KV_LEVELS; z++) {
    const int k = k0 + z;
    if (k >= L1) continue;
    double sum = 0.0;
    int cnt = 0;
#pragma unroll
    for (int l = 0; l < 4; l++) {
      if (l >= N || fkt[l] == 0) continue;
      const int kk = k < fkb[l] ? fkb[l] : (k > fkt[l] ? fkt[l] : k);
      sum = sum + A0(w.turb_kvf, kk, fc[l], J);
      cnt++;
    }
    A0(s.Kv, k, i, E) = cnt == 0 ? (double)1.0e-2f : sum / cnt;
  }
}
