// k_hydro.hip - hydrodynamic stages of the step: bottom friction, ELM advection,
// momentum column solves, free-surface assembly, velocities, layers, node velocities.
//
// Each kernel restates one reference loop (file:line cited per kernel, paths relative to
// GOM_Rv1.1.0/source) in the same floating-point evaluation order, so that with
// contraction disabled (-ffp-contract=off) the FP64 results are bit-identical to the
// reference's serial run on the same inputs. One thread owns one face, cell or node column
// (column recurrences) or one (layer, entity) pair (independent updates).
#include "gomgpu_dev.h"

// ================================================================ bottom friction
// calculate_bottom_friction.f90:21-62. bf_flag 1 uses the Manning drag of the still-water
// face depth, computed once on the host with the same libm pow as the reference.
__global__ void k_friction(DevMesh m, DevState s, DevDims d, int bf_flag, double dry, double vk) {
  int j = gom_block() * blockDim.x + threadIdx.x;
  if (j >= d.J) return;
  const int J = d.J;
  double cdb = 0.0, gb = 0.0;
  if (bf_flag != 0 && m.h_face[j] > dry) {
    int k = m.bottom_layer_at_face[j];
    if (bf_flag == 1) {
      cdb = m.Cdb_static[j];
    } else {
      // the reference's host log (correctly rounded), not the device library's
      double lg = gom_log(dmaxf(0.5, A0(s.dzhalf_face, k - 1, j, J)) / m.bf_height[j]);
      cdb = vk * vk / (lg * lg);
    }
    double u = AL(s.un_face, k, j, J), v = AL(s.vn_face, k, j, J);
    gb = cdb * sqrt(u * u + v * v);
  }
  s.Cdb[j] = cdb;
  s.Gamma_B[j] = gb;
}

// gom_log on the host, for checking it against the C library (test use).
extern "C" void gom_log_array(const double *x, double *y, long n) {
  for (long i = 0; i < n; i++) y[i] = gom_log(x[i]);
}
// gom_cos over a host array (the host build of the device cosine, for its tests)
extern "C" void gom_cos_array(const double *x, double *y, long n) {
  for (long i = 0; i < n; i++) y[i] = gom_cos(x[i]);
}

// ================================================================ ELM
// Sub-step counts per node-layer: solve_nonlinear_advection.f90:52-87
__global__ void k_elm_substeps(DevMesh m, DevState s, DevDims d, double dt, double dzmin,
                               int imin, int imax) {
  int t = gom_block() * blockDim.x + threadIdx.x;
  if (t >= d.P * d.L) return;
  const int P = d.P, E = d.E;
  int i = t % P, k = t / P + 1;
  int nadj = m.adj_cells_at_node[i];
  double summ = 0.0;
  double ui = (A0(s.u_node, k, i, P) + A0(s.u_node, k - 1, i, P)) * 0.5;
  double vi = (A0(s.v_node, k, i, P) + A0(s.v_node, k - 1, i, P)) * 0.5;
  double xi = m.x_node[i], yi = m.y_node[i];
  for (int jj = 1; jj <= nadj; jj++) {
    int ie = TB(m.adj_cellnum_at_node, jj, i, P) - 1;
    int id = TB(m.node_count_each_element, jj, i, P);
    int N = m.tri_or_quad[ie];
    double devm = 0.0;
    for (int l = 1; l <= 2; l++) {
      int nd = (l == 1) ? TB(m.nodenum_at_cell, sen(N, id, 1), ie, E) - 1
                        : TB(m.nodenum_at_cell, sen(N, id, N - 1), ie, E) - 1;
      double dx = xi - m.x_node[nd], dy = yi - m.y_node[nd];
      double rl = sqrt(dx * dx + dy * dy);
      double u2 = (A0(s.u_node, k, nd, P) + A0(s.u_node, k - 1, nd, P)) * 0.5;
      double v2 = (A0(s.v_node, k, nd, P) + A0(s.v_node, k - 1, nd, P)) * 0.5;
      double du = ui - u2, dv = vi - v2;
      double dev = sqrt(du * du + dv * dv) / rl * dt;
      if (dev > devm) devm = dev;
    }
    double vw = (A0(s.w_node, k, i, P) + A0(s.w_node, k - 1, i, P)) * 0.5;
    int iw = (int)(2 * vw * dt / dzmin);
    int i1 = (int)(devm / (double)1.e-1f);
    int idelta = i1 > iw ? i1 : iw;
    summ = summ + (double)idelta / nadj;
  }
  int ns = (int)summ;
  ns = ns < imax ? ns : imax;
  ns = ns > imin ? ns : imin;
  AL(s.num_sub_elm_iteration, k, i, P) = ns;
}

// check_intersection.f90:14-50
__device__ __forceinline__ int check_isect(double x1, double x2, double x3, double x4, double y1,
                                           double y2, double y3, double y4, double &xin, double &yin) {
  double delta = (x2 - x1) * (y3 - y4) - (y2 - y1) * (x3 - x4);
  double delta1 = (x3 - x1) * (y3 - y4) - (y3 - y1) * (x3 - x4);
  double delta2 = (x2 - x1) * (y3 - y1) - (y2 - y1) * (x3 - x1);
  if (delta != 0.0) {
    double tt1 = delta1 / delta, tt2 = delta2 / delta;
    if (tt1 >= -0.0 && tt1 <= 1.0 && tt2 >= -0.0 && tt2 <= 1.0) {
      xin = x1 + (x2 - x1) * tt1;
      yin = y1 + (y2 - y1) * tt1;
      return 1;
    }
  }
  return 0;
}

// bilinear_interpolation.f90:8-121; returns 0 or GOMGPU_ERR_BILINEAR
__device__ __forceinline__ int bilinear(double x1, double x2, double x3, double x4, double y1, double y2, double y3,
                        double y4, double x, double y, double *shapef) {
  const double small3 = (double)1.e-5f;
  const double lim = (double)1.1f;
  double x0 = (x1 + x2 + x3 + x4) / 4.0, y0 = (y1 + y2 + y3 + y4) / 4.0;
  double axi0 = x2 - x1 + x3 - x4, axi1 = y2 - y1 + y3 - y4;
  double aet0 = x3 + x4 - x1 - x2, aet1 = y3 + y4 - y1 - y2;
  double bxy0 = x1 - x2 + x3 - x4, bxy1 = y1 - y2 + y3 - y4;
  double dxi = 2.0 * ((x3 - x4) * (y1 - y2) - (y3 - y4) * (x1 - x2));
  double deta = 2.0 * ((x4 - x1) * (y3 - y2) - (y4 - y1) * (x3 - x2));
  double xi = 0.0, eta = 0.0, dd;
  if ((fabs(bxy0) < small3 && fabs(bxy1) < small3) || (fabs(dxi) < small3 && fabs(deta) < small3)) {
    dd = axi0 * aet1 - axi1 * aet0;
    if (dd == 0.0) return GOMGPU_ERR_BILINEAR;
    xi = 4.0 * (aet1 * (x - x0) - aet0 * (y - y0)) / dd;
    eta = 4.0 * (axi0 * (y - y0) - axi1 * (x - x0)) / dd;
  } else if (fabs(dxi) < small3 && fabs(deta) >= small3) {
    eta = 4.0 * (bxy1 * (x - x0) - bxy0 * (y - y0)) / deta;
    double t1 = axi0 + eta * bxy0, t2 = axi1 + eta * bxy1;
    dd = t1 * t1 + t2 * t2;
    if (dd == 0.0) return GOMGPU_ERR_BILINEAR;
    xi = ((4.0 * (x - x0) - eta * aet0) * (axi0 + eta * bxy0) + (4.0 * (y - y0) - eta * aet1) * (axi1 + eta * bxy1)) / dd;
  } else if (fabs(dxi) >= small3 && fabs(deta) < small3) {
    xi = 4.0 * (bxy1 * (x - x0) - bxy0 * (y - y0)) / dxi;
    double t1 = aet0 + xi * bxy0, t2 = aet1 + xi * bxy1;
    dd = t1 * t1 + t2 * t2;
    if (dd == 0.0) return GOMGPU_ERR_BILINEAR;
    eta = ((4.0 * (x - x0) - xi * axi0) * (aet0 + xi * bxy0) + (4.0 * (y - y0) - xi * axi1) * (aet1 + xi * bxy1)) / dd;
  } else {
    double beta = aet1 * axi0 - aet0 * axi1 - 4 * (bxy1 * (x - x0) - bxy0 * (y - y0));
    double gamma = 4.0 * (aet0 * (y - y0) - aet1 * (x - x0));
    double delta = beta * beta - 4 * gamma * dxi;
    if (delta == 0.0) {
      xi = -beta / 2 / dxi;
      eta = (4.0 * (bxy1 * (x - x0) - bxy0 * (y - y0)) - xi * dxi) / deta;
    } else if (delta > 0) {
      double rx[2], re[2];
      rx[0] = (-beta + sqrt(delta)) / 2 / dxi;
      rx[1] = (-beta - sqrt(delta)) / 2 / dxi;
      int icount = 0;
      for (int i = 0; i < 2; i++) {
        re[i] = (4 * (bxy1 * (x - x0) - bxy0 * (y - y0)) - rx[i] * dxi) / deta;
        if (fabs(rx[i]) <= lim && fabs(re[i]) <= lim) {
          xi = rx[i];
          eta = re[i];
          icount++;
        }
      }
      if (icount == 2 && fabs(rx[0] - rx[1]) < (double)1.e-6f) {
      } else if (icount != 1) {
        return GOMGPU_ERR_BILINEAR;
      }
    } else {
      return GOMGPU_ERR_BILINEAR;
    }
  }
  if (fabs(xi) > lim || fabs(eta) > lim) return GOMGPU_ERR_BILINEAR;
  xi = dminf(1.0, dmaxf(xi, -1.0));
  eta = dminf(1.0, dmaxf(eta, -1.0));
  shapef[0] = (1 - xi) * (1 - eta) / 4;
  shapef[1] = (1 + xi) * (1 - eta) / 4;
  shapef[2] = (1 + xi) * (1 + eta) / 4;
  shapef[3] = (1 - xi) * (1 + eta) / 4;
  return 0;
}

struct ElmCtx {
  DevMesh m;
  DevState s;
  DevDims d;
  double MSL;
};

// One visited cell of the walk, read from the cell records with a few aligned loads issued
// together: node count, area, node coordinates, node ids and neighbours in slot order (the
// values of tri_or_quad, area, x_node/y_node, nodenum_at_cell and adj_cellnum_at_cell).
struct CellRec {
  int N;
  double ar;
  double X[4], Y[4];
  int nd[4], adj[4];
};

__device__ __forceinline__ void cell_rec(const DevMesh &m, int e, CellRec &r) {
  r.N = m.tri_or_quad[e];
  r.ar = m.area[e];
  const double2 *px = (const double2 *)(m.cg_x + 4 * (size_t)e), *py = (const double2 *)(m.cg_y + 4 * (size_t)e);
  const double2 x01 = px[0], x23 = px[1], y01 = py[0], y23 = py[1];
  const int4 nn = *(const int4 *)(m.cg_node + 4 * (size_t)e), aa = *(const int4 *)(m.cg_adj + 4 * (size_t)e);
  r.X[0] = x01.x, r.X[1] = x01.y, r.X[2] = x23.x, r.X[3] = x23.y;
  r.Y[0] = y01.x, r.Y[1] = y01.y, r.Y[2] = y23.x, r.Y[3] = y23.y;
  r.nd[0] = nn.x, r.nd[1] = nn.y, r.nd[2] = nn.z, r.nd[3] = nn.w;
  r.adj[0] = aa.x, r.adj[1] = aa.y, r.adj[2] = aa.z, r.adj[3] = aa.w;
}

// Element i (0..3) of a register-held 4-array: a select chain, so the array stays in
// registers (a dynamic index would move it to scratch memory).
template <typename T>
__device__ __forceinline__ T pick4(const T (&a)[4], int i) {
  return i == 0 ? a[0] : i == 1 ? a[1] : i == 2 ? a[2] : a[3];
}

// Sum over the cell's edges l = 1..N of |area(node l, node sen(l,1), point)|, in slot order
// (the containment test of search_straight_line.f90:60-100). Unrolled over 4 slots; the
// terms of slots beyond N are not added.
__device__ __forceinline__ double area_sum(const CellRec &g, double px, double py) {
  double sum = 0.0;
#pragma unroll
  for (int l = 1; l <= 4; l++) {
    const double xb = l < g.N ? g.X[l & 3] : g.X[0], yb = l < g.N ? g.Y[l & 3] : g.Y[0];
    const double t = sum + fabs(carea(g.X[l - 1], xb, px, g.Y[l - 1], yb, py));
    sum = l <= g.N ? t : sum;
  }
  return sum;
}

// Whether (px, py) lies within 1e-12 of an edge of the cell, relative to the edge's length:
// |(B - A) x (P - A)| <= 1e-12 |B - A|^2 for some edge A -> B. The containment and crossing
// tests (search_straight_line.f90:60-100, check_intersection.f90:34-48, inclusive bounds)
// can assign such a point to either cell sharing the edge on the last bit of its
// coordinates, so its departure cell is reproducible only by the same arithmetic.
__device__ __forceinline__ bool near_edge(const CellRec &g, double px, double py) {
  bool near = false;
#pragma unroll
  for (int l = 0; l < 4; l++) {
    const double xb = l + 1 < g.N ? g.X[(l + 1) & 3] : g.X[0], yb = l + 1 < g.N ? g.Y[(l + 1) & 3] : g.Y[0];
    const double ex = xb - g.X[l], ey = yb - g.Y[l];
    const double cr = ex * (py - g.Y[l]) - ey * (px - g.X[l]);
    near = near || (l < g.N && fabs(cr) <= 1.0e-12 * (ex * ex + ey * ey));
  }
  return near;
}

// The first edge l (slots sen(l,1), sen(l,2), l = 1..N in order) that the segment
// (xa,ya)-(xb,yb) meets, skipping the edge with node ids {md1, md2} when md1 > 0
// (search_straight_line.f90:144-169, 320-350). Every edge is tested; the lowest hit wins,
// which is the edge the reference's loop stops at. Returns l, or 0 if none; (xin, yin) is
// the intersection point of that edge.
__device__ __forceinline__ int first_edge(const CellRec &g, double xa, double xb, double ya, double yb, int md1,
                                          int md2, double &xin, double &yin) {
  int hit_l = 0;
#pragma unroll
  for (int l = 4; l >= 1; l--) {
    const int i1 = l < g.N ? l : 0;                                  // sen(N, l, 1) - 1
    const int i2 = l + 2 > g.N ? l + 1 - g.N : l + 1;                // sen(N, l, 2) - 1
    const int jd1 = pick4(g.nd, i1), jd2 = pick4(g.nd, i2);
    const bool skip = md1 > 0 && ((jd1 == md1 && jd2 == md2) || (jd2 == md1 && jd1 == md2));
    double xi = 0.0, yi = 0.0;
    const int hit = check_isect(xa, xb, pick4(g.X, i1), pick4(g.X, i2), ya, yb, pick4(g.Y, i1), pick4(g.Y, i2), xi, yi);
    if (l <= g.N && !skip && hit == 1) {
      hit_l = l;
      xin = xi;
      yin = yi;
    }
  }
  return hit_l;
}

// search_straight_line.f90:15-399 (mode 1): straight-line walk from (x0,y0) to (xt,yt).
// On return rc holds the record of the end cell nnel.
__device__ __forceinline__ int search_line(const ElmCtx &c, int &nnel, int &jlev, double time, double x0, double y0,
                                           double z0, double &xt, double &yt, double &zt, int &nfl, CellRec &rc) {
  const DevMesh &m = c.m;
  const DevState &s = c.s;
  const int E = c.d.E, J = c.d.J, L = c.d.L;
  const double small06 = (double)1.e-6f, MSL = c.MSL;
  int nel = nnel;  // 1-based
  int e0 = nel - 1;
  cell_rec(m, e0, rc);
  if (s.top_layer_at_element[e0] == 0 || MSL - m.h_cell[e0] > MSL + s.eta_cell[e0]) return GOMGPU_ERR_ELM_DRY_START;
  nfl = 0;
  double trm = time;
  double area1 = area_sum(rc, x0, y0), area2 = area_sum(rc, xt, yt);
  double ae = fabs(area1 - rc.ar) / rc.ar;
  if (ae > small06) return GOMGPU_ERR_ELM_NOT_IN_CELL;
  ae = fabs(area2 - rc.ar) / rc.ar;
  if (ae < small06) {
    nnel = nel;
  } else {
    double xcg = (1.0 - 1.0e-4) * x0 + 1.0e-4 * m.x_cell[e0];
    double ycg = (1.0 - 1.0e-4) * y0 + 1.0e-4 * m.y_cell[e0];
    double pathl = sqrt((xt - xcg) * (xt - xcg) + (yt - ycg) * (yt - ycg));
    if ((xcg == xt && ycg == yt) || pathl == 0.0) return GOMGPU_ERR_ELM_ZERO_PATH;
    double xin = 0.0, yin = 0.0;
    int nel_j = first_edge(rc, xcg, xt, ycg, yt, 0, 0, xin, yin);
    if (nel_j == 0) return GOMGPU_ERR_ELM_NO_EDGE;
    double zin = z0;
    int iteration = 0;
    for (;;) {
      iteration++;
      if (iteration > 1000) {
        nfl = 1;
        xt = xin;
        yt = yin;
        zt = zin;
        nnel = nel;
        break;
      }
      e0 = nel - 1;
      const int N = rc.N;
      const int md1 = pick4(rc.nd, sen(N, nel_j, 1) - 1);
      const int md2 = pick4(rc.nd, sen(N, nel_j, 2) - 1);
      const int adj = pick4(rc.adj, nel_j - 1);
      // the neighbour's record and wet flag, loaded together (row 1 stands in for none)
      CellRec nx;
      cell_rec(m, adj > 0 ? adj - 1 : 0, nx);
      const int adj_top = s.top_layer_at_element[adj > 0 ? adj - 1 : 0];
      double dist = sqrt((xin - xt) * (xin - xt) + (yin - yt) * (yin - yt));
      if (dist / pathl > 1.0 + 1.0e-4) return GOMGPU_ERR_ELM_OVERSHOT;
      zin = zt - dist / pathl * (zt - zin);
      trm = trm * dist / pathl;
      pathl = sqrt((xin - xt) * (xin - xt) + (yin - yt) * (yin - yt));
      if (pathl == 0.0 || trm == 0.0) return GOMGPU_ERR_ELM_TARGET;
      if (adj == 0 || adj_top == 0) {
        int isd = TB(m.facenum_at_cell, nel_j, e0, E) - 1;
        if (TB(m.nodenum_at_face, 1, isd, J) + TB(m.nodenum_at_face, 2, isd, J) != md1 + md2)
          return GOMGPU_ERR_ELM_WRONG_SIDE;
        xin = (1.0 - 1.0e-4) * xin + 1.0e-4 * m.x_cell[e0];
        yin = (1.0 - 1.0e-4) * yin + 1.0e-4 * m.y_cell[e0];
        // The reference computes a sliding target here and then ends the walk at the
        // nudged boundary point (search_straight_line.f90:246-282).
        nfl = 1;
        xt = xin;
        yt = yin;
        zt = zin;
        nnel = nel;
        break;
      }
      nel = adj;
      e0 = nel - 1;
      rc = nx;
      area1 = area_sum(rc, xt, yt);
      ae = fabs(area1 - rc.ar) / rc.ar;
      if (ae < small06) {
        nnel = nel;
        break;
      }
      nel_j = first_edge(rc, xcg, xt, ycg, yt, md1, md2, xin, yin);
      if (nel_j == 0) return GOMGPU_ERR_ELM_NO_EDGE;
    }
  }
  e0 = nnel - 1;
  if (s.top_layer_at_element[e0] == 0 || MSL - m.h_cell[e0] > MSL + s.eta_cell[e0]) return GOMGPU_ERR_ELM_DRY_END;
  double z = dminf(dmaxf(zt, MSL - m.h_cell[e0]), MSL + s.eta_cell[e0]);
  zt = z;
  if (z <= m.z_level[0] || z > m.z_level[L]) return GOMGPU_ERR_ELM_VERTICAL;
  // the level k with z_level[k] < z <= z_level[k+1] (z_level increases strictly): bisection
  int lo = 0, hi = L;  // z_level[lo] < z <= z_level[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (z > m.z_level[mid])
      lo = mid;
    else
      hi = mid;
  }
  jlev = lo + 1;
  int bl = m.bottom_layer_at_element[e0], tl = s.top_layer_at_element[e0];
  int jl = jlev < bl ? bl : jlev;
  jlev = jl > tl ? tl : jl;
  return 0;
}

// The end of a backtrace: the foot point, the top of its layer and the foot's interpolation
// weights in the cell that holds it (the inputs of the tracer interpolation).
struct ElmFoot {
  double xt, yt, zt, zup, st[4];
};

// ELM_backtrace_v1.f90:15-200 (i_which_backtrack = 1). near is set when the end point of
// any sub-step lies within 1e-12 of an edge of its cell (near_edge).
__device__ __forceinline__ int backtrace(const ElmCtx &c, int bt_step, double bt_dt, double &uu, double &vv, double &ww,
                         double x0, double y0, double z0, int &nnel, int &jlev, ElmFoot &ft, bool &near) {
  const DevMesh &m = c.m;
  const DevState &s = c.s;
  const int E = c.d.E, P = c.d.P;
  double xt = 0, yt = 0, zt = 0;
  for (int idt = 1; idt <= bt_step; idt++) {
    xt = x0 - bt_dt * uu;
    yt = y0 - bt_dt * vv;
    zt = z0 - bt_dt * ww;
    int iflqs1 = 0;
    CellRec rc;
    int err = search_line(c, nnel, jlev, bt_dt, x0, y0, z0, xt, yt, zt, iflqs1, rc);
    if (err) return err;
    near = near || near_edge(rc, xt, yt);
    int e0 = nnel - 1, N = rc.N, jl = jlev;
    double zup = (jl == s.top_layer_at_element[e0]) ? c.MSL + s.eta_cell[e0] : m.z_level[jl];
    if (zt > zup) return GOMGPU_ERR_ELM_VERTICAL;
    double zrat = (zup - zt) / AL(s.dz_cell, jl, e0, E);
    double vxn[4], vyn[4], vzn[4], st[4] = {0.0, 0.0, 0.0, 0.0};
    int nd[4];
    for (int l = 0; l < N; l++) {
      nd[l] = pick4(rc.nd, l) - 1;
      double ux1 = A0(s.u_node, jl, nd[l], P), ux0 = A0(s.u_node, jl - 1, nd[l], P);
      double uy1 = A0(s.v_node, jl, nd[l], P), uy0 = A0(s.v_node, jl - 1, nd[l], P);
      double uz1 = A0(s.w_node, jl, nd[l], P), uz0 = A0(s.w_node, jl - 1, nd[l], P);
      vxn[l] = ux1 * (1 - zrat) + ux0 * zrat;
      vyn[l] = uy1 * (1 - zrat) + uy0 * zrat;
      vzn[l] = uz1 * (1 - zrat) + uz0 * zrat;
    }
    double X1 = rc.X[0], X2 = rc.X[1], X3 = rc.X[2];
    double Y1 = rc.Y[0], Y2 = rc.Y[1], Y3 = rc.Y[2];
    double ar = rc.ar;
    if (N == 3) {
      st[0] = carea(xt, X2, X3, yt, Y2, Y3) / ar;
      st[1] = carea(X1, xt, X3, Y1, yt, Y3) / ar;
      st[0] = dmaxf(0.0, dminf(1.0, st[0]));
      st[1] = dmaxf(0.0, dminf(1.0, st[1]));
      if (st[0] + st[1] > 1) {
        st[2] = 0;
        st[1] = 1.0 - st[0];
      } else {
        st[2] = 1.0 - st[0] - st[1];
      }
    } else {
      double X4 = rc.X[3], Y4 = rc.Y[3];
      double aa1 = carea(X1, X2, xt, Y1, Y2, yt), aa2 = carea(X2, X3, xt, Y2, Y3, yt);
      double aa3 = carea(X3, X4, xt, Y3, Y4, yt), aa4 = carea(X4, X1, xt, Y4, Y1, yt);
      double aa = fabs(aa1) + fabs(aa2) + fabs(aa3) + fabs(aa4);
      if (fabs(aa - ar) / ar > (double)1.e-6f) return GOMGPU_ERR_ELM_OUTSIDE_QUAD;
      int e2 = bilinear(X1, X2, X3, X4, Y1, Y2, Y3, Y4, xt, yt, st);
      if (e2) return e2;
    }
    uu = 0.0;
    vv = 0.0;
    ww = 0.0;
    for (int l = 0; l < N; l++) {
      uu = uu + vxn[l] * st[l];
      vv = vv + vyn[l] * st[l];
      ww = ww + vzn[l] * st[l];
    }
    ft.xt = xt;
    ft.yt = yt;
    ft.zt = zt;
    ft.zup = zup;
#pragma unroll
    for (int l = 0; l < 4; l++) ft.st[l] = st[l];
    if (iflqs1 == 1) break;
    x0 = xt;
    y0 = yt;
    z0 = zt;
  }
  return 0;
}

// solve_nonlinear_advection.f90:92-202; one thread per (layer, face). Departures within 1e-12
// of a cell edge are counted in near_count. With dep non-NULL, the departure cell (negated
// for such a point) and level of face-layer t = (k-1) J + j go to dep[t] and dep[L J + t]
// (0 where nothing is backtraced).
__global__ void __launch_bounds__(256) GOM_WAVES(GOM_WAVES_ELM) k_elm(DevMesh m, DevState s, DevDims d, double MSL, double dt, double onoff,
                                             int *err, int32_t *dep, unsigned long long *near_count) {
  const int J = d.J, E = d.E, P = d.P;
  int j, k;
  if (!gom_tile(J, d.L, k, j)) return;
  const size_t t = (size_t)k * J + j;  // face-layer (k+1, j)
  k += 1;
  int kb = m.bottom_layer_at_face[j], kt = s.top_layer_at_face[j];
  double un = AL(s.un_face, k, j, J), vn = AL(s.vn_face, k, j, J);
  if (dep) {
    dep[t] = 0;
    dep[(size_t)J * d.L + t] = 0;
  }
  if (k < kb || k > kt) {
    AL(s.un_ELM, k, j, J) = 0.0;
    AL(s.vn_ELM, k, j, J) = 0.0;
    return;
  }
  int n1 = TB(m.nodenum_at_face, 1, j, J) - 1, n2 = TB(m.nodenum_at_face, 2, j, J) - 1;
  double atod = dmaxf(m.h_node[n1], m.h_node[n2]);
  int a1 = TB(m.adj_cellnum_at_face, 1, j, J), a2 = TB(m.adj_cellnum_at_face, 2, j, J);
  int ie0 = 0;
  if (!(fabs(atod) < onoff)) {
    if (s.top_layer_at_element[a1 - 1] != 0)
      ie0 = a1;
    else if (a2 != 0 && s.top_layer_at_element[a2 - 1] != 0)
      ie0 = a2;
  }
  if (ie0 == 0) {
    AL(s.un_ELM, k, j, J) = un;
    AL(s.vn_ELM, k, j, J) = vn;
    return;
  }
  double ct = m.cos_theta[j], st = m.sin_theta[j];
  double z0;
  if (k == s.top_layer_at_element[ie0 - 1])
    z0 = MSL + s.eta_cell[ie0 - 1] - AL(s.dz_cell, k, ie0 - 1, E) * 0.5;
  else
    z0 = m.z_level[k] - AL(s.dz_cell, k, ie0 - 1, E) * 0.5;
  double uu = un * ct - vn * st;
  double vv = un * st + vn * ct;
  double wdown = (A0(s.w_node, k - 1, n1, P) + A0(s.w_node, k - 1, n2, P)) * 0.5;
  double wup = (A0(s.w_node, k, n1, P) + A0(s.w_node, k, n2, P)) * 0.5;
  double ww = (wdown + wup) * 0.5;
  double vmag = sqrt(uu * uu + vv * vv + ww * ww);
  if (vmag <= (double)1.e-4f) {
    AL(s.un_ELM, k, j, J) = un;
    AL(s.vn_ELM, k, j, J) = vn;
    return;
  }
  int bt_step = (int)((AL(s.num_sub_elm_iteration, k, n1, P) + AL(s.num_sub_elm_iteration, k, n2, P)) * 0.5);
  double bt_dt = dt / bt_step;
  ElmCtx c{m, s, d, MSL};
  int nnel = ie0, jlev = k;
  ElmFoot ft;
  bool near = false;
  int e = backtrace(c, bt_step, bt_dt, uu, vv, ww, m.x_face[j], m.y_face[j], z0, nnel, jlev, ft, near);
  if (e) latch(err, e, j + 1);
  if (near) atomicAdd(near_count, 1ull);
  if (dep) {
    dep[t] = near ? -nnel : nnel;
    dep[(size_t)J * d.L + t] = jlev;
  }
  AL(s.un_ELM, k, j, J) = uu * ct + vv * st;
  AL(s.vn_ELM, k, j, J) = -uu * st + vv * ct;
}

// main.f90:171-176 (advection off)
__global__ void k_elm_copy(DevState s, long n) {
  long t = gom_block() * (long)blockDim.x + threadIdx.x;
  if (t >= n) return;
  s.un_ELM[t] = s.un_face[t];
  s.vn_ELM[t] = s.vn_face[t];
}

// ================================================================ ELM scalar transport
// solve_transport_equation_ELM_v8.f90:7-346 (transport_solver 2). Salt and temperature of
// each wet cell-layer are backtraced from the layer centre, take the surface heat source and
// an explicit horizontal diffusion term, then vertical diffusion implicit per column; the
// node values are area*dz weighted cell means, the face values node means.

// Node (or face) value blended between layers k1 and k2 of an (L, n) field.
__device__ __forceinline__ double zblend(const double *a, int k1, int k2, double zrat2, int e, int n) {
  return AL(a, k1, e, n) * (1 - zrat2) + AL(a, k2, e, n) * zrat2;
}

// The tracers at a backtraced foot (ELM_backtrace_v1.f90:216-473, which = 2). A boundary
// cell gives its boundary values; otherwise the node values at the two levels bracketing zt
// are weighted by the foot's interpolation weights (or averaged over the wet nodes), and a
// wet sub-element holding the foot -- for a triangle the four sub-triangles through the face
// midpoints, for a quad the four sub-quads through the face midpoints and the centre --
// replaces that by its own interpolation. Boundary values are sub-cycle 0 of con_at_ob /
// con_at_Qb, salt in tracer slot 1 and temperature in slot 2. nnel is 1-based.
__device__ __forceinline__ int elm_tracer_foot(const DevMesh &m, const DevState &s, const DevForcing &f,
                                               const DevDims &d, int Nt, int nnel, int jlev, const ElmFoot &ft,
                                               double &tt, double &ss) {
  const int E = d.E, J = d.J, P = d.P, NT = d.NT;
  const int e0 = nnel - 1;
  const double small_06 = (double)1.e-6f;
  tt = 0.0;
  ss = 0.0;
  const int ib1 = m.ob_element_flag[e0], ib2 = m.Qb_element_flag[e0];
  if (ib1 > 0) {
    ss = f.con_at_ob[(size_t)Nt * (0 + (size_t)NT * (ib1 - 1))];
    tt = f.con_at_ob[(size_t)Nt * (1 + (size_t)NT * (ib1 - 1))];
    return 0;
  }
  if (ib2 > 0) {
    ss = f.con_at_Qb[(size_t)Nt * (0 + (size_t)NT * (ib2 - 1))];
    tt = f.con_at_Qb[(size_t)Nt * (1 + (size_t)NT * (ib2 - 1))];
    return 0;
  }
  const double xt = ft.xt, yt = ft.yt, zt = ft.zt;
  const double zanchor = ft.zup - AL(s.dz_cell, jlev, e0, E) * 0.5;
  const int bl = m.bottom_layer_at_element[e0];
  int k1, k2;
  double zrat2;
  if (jlev == s.top_layer_at_element[e0] && zt > zanchor) {
    k1 = jlev;
    k2 = jlev;
    zrat2 = 0.5;
  } else if (zt > zanchor) {
    k1 = jlev;
    k2 = jlev + 1;
    zrat2 = (zt - zanchor) / A0(s.dzhalf_cell, jlev, e0, E);
  } else {
    k1 = jlev - 1 > bl ? jlev - 1 : bl;
    k2 = jlev;
    zrat2 = 1.0 - (zanchor - zt) / A0(s.dzhalf_cell, k1, e0, E);
  }
  zrat2 = dmaxf(0.0, dminf(1.0, zrat2));
  const int N = m.tri_or_quad[e0];
  int icount = 0;
  double ta = 0.0, sa = 0.0, t_xi[4] = {0.0, 0.0, 0.0, 0.0}, s_xi[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int l = 0; l < 4; l++) {
    if (l < N) {
      const int nd = TB(m.nodenum_at_cell, l + 1, e0, E) - 1;
      if (s.top_layer_at_node[nd] != 0) {
        icount++;
        t_xi[l] = zblend(s.temp_node, k1, k2, zrat2, nd, P);
        s_xi[l] = zblend(s.salt_node, k1, k2, zrat2, nd, P);
        ta = ta + t_xi[l];
        sa = sa + s_xi[l];
      }
    }
  }
  if (icount == 0) return GOMGPU_ERR_ELM_ALL_DRY;
  ta = ta / icount;
  sa = sa / icount;
  if (icount != N) {
    tt = ta;
    ss = sa;
  } else {
#pragma unroll
    for (int l = 0; l < 4; l++) {
      if (l < N) {
        tt = tt + t_xi[l] * ft.st[l];
        ss = ss + s_xi[l] * ft.st[l];
      }
    }
  }
  int found = 0;
  const double ar = m.area[e0];
  if (N == 3) {
    for (int l = 1; l <= 4; l++) {
      int n1, f2, f3;
      double X1, Y1;
      if (l <= 3) {
        n1 = TB(m.nodenum_at_cell, l, e0, E);
        f2 = TB(m.facenum_at_cell, sen(N, l, 2), e0, E);
        f3 = TB(m.facenum_at_cell, sen(N, l, 1), e0, E);
        X1 = m.x_node[n1 - 1];
        Y1 = m.y_node[n1 - 1];
      } else {
        n1 = TB(m.facenum_at_cell, 1, e0, E);  // the centre sub-triangle: a face midpoint first
        f2 = TB(m.facenum_at_cell, 2, e0, E);
        f3 = TB(m.facenum_at_cell, 3, e0, E);
        X1 = m.x_face[n1 - 1];
        Y1 = m.y_face[n1 - 1];
      }
      const double X2 = m.x_face[f2 - 1], Y2 = m.y_face[f2 - 1], X3 = m.x_face[f3 - 1], Y3 = m.y_face[f3 - 1];
      const double aa1 = carea(xt, X2, X3, yt, Y2, Y3);
      const double aa2 = carea(X1, xt, X3, Y1, yt, Y3);
      const double aa3 = carea(X1, X2, xt, Y1, Y2, yt);
      const double aa = fabs(aa1) + fabs(aa2) + fabs(aa3);
      const double subrat = fabs(aa - ar / 4) * 4 / ar;
      if (subrat < 100 * small_06) {
        found = 1;
        double sig0 = aa1 * 4 / ar, sig1 = aa2 * 4 / ar, sig2;
        sig0 = dmaxf(0.0, dminf(1.0, sig0));
        sig1 = dmaxf(0.0, dminf(1.0, sig1));
        if (sig0 + sig1 > 1) {
          sig2 = 0;
          sig1 = 1 - sig0;
        } else {
          sig2 = 1 - sig0 - sig1;
        }
        bool wet = l <= 3 ? s.top_layer_at_node[n1 - 1] != 0 : s.top_layer_at_face[n1 - 1] != 0;
        wet = wet && s.top_layer_at_face[f2 - 1] != 0 && s.top_layer_at_face[f3 - 1] != 0;
        if (wet) {
          const double t0 = l <= 3 ? zblend(s.temp_node, k1, k2, zrat2, n1 - 1, P)
                                   : zblend(s.temp_face, k1, k2, zrat2, n1 - 1, J);
          const double s0 = l <= 3 ? zblend(s.salt_node, k1, k2, zrat2, n1 - 1, P)
                                   : zblend(s.salt_face, k1, k2, zrat2, n1 - 1, J);
          const double t1 = zblend(s.temp_face, k1, k2, zrat2, f2 - 1, J);
          const double t2 = zblend(s.temp_face, k1, k2, zrat2, f3 - 1, J);
          const double s1 = zblend(s.salt_face, k1, k2, zrat2, f2 - 1, J);
          const double s2 = zblend(s.salt_face, k1, k2, zrat2, f3 - 1, J);
          tt = t0 * sig0 + t1 * sig1 + t2 * sig2;
          ss = s0 * sig0 + s1 * sig1 + s2 * sig2;
          break;
        }
      }
    }
  } else {
    const double XC = m.x_cell[e0], YC = m.y_cell[e0];
    for (int l = 1; l <= 4; l++) {
      const int n1 = TB(m.nodenum_at_cell, l, e0, E);
      const int f2 = TB(m.facenum_at_cell, sen(N, l, 3), e0, E);
      const int f4 = TB(m.facenum_at_cell, sen(N, l, 2), e0, E);
      const double X1 = m.x_node[n1 - 1], Y1 = m.y_node[n1 - 1];
      const double X2 = m.x_face[f2 - 1], Y2 = m.y_face[f2 - 1];
      const double X4 = m.x_face[f4 - 1], Y4 = m.y_face[f4 - 1];
      const double aa1 = carea(X1, X2, xt, Y1, Y2, yt);
      const double aa2 = carea(X2, XC, xt, Y2, YC, yt);
      const double aa3 = carea(XC, X4, xt, YC, Y4, yt);
      const double aa4 = carea(X4, X1, xt, Y4, Y1, yt);
      const double bb1 = carea(X1, X2, X4, Y1, Y2, Y4);
      const double bb2 = carea(X2, XC, X4, Y2, YC, Y4);
      if (bb1 <= 0 || bb2 <= 0) return GOMGPU_ERR_ELM_NEG_SUBCELL;
      const double bb = bb1 + bb2;
      const double aa = fabs(aa1) + fabs(aa2) + fabs(aa3) + fabs(aa4);
      const double subrat = fabs(aa - bb) / bb;
      if (subrat < 4.0 * small_06) {
        found = 1;
        double sq[4];
        const int e2 = bilinear(X1, X2, XC, X4, Y1, Y2, YC, Y4, xt, yt, sq);
        if (e2) return e2;
        if (s.top_layer_at_node[n1 - 1] != 0 && s.top_layer_at_face[f2 - 1] != 0 &&
            s.top_layer_at_face[f4 - 1] != 0) {
          const double t0 = zblend(s.temp_node, k1, k2, zrat2, n1 - 1, P);
          const double t1 = zblend(s.temp_face, k1, k2, zrat2, f2 - 1, J);
          const double t3 = zblend(s.temp_face, k1, k2, zrat2, f4 - 1, J);
          const double s0 = zblend(s.salt_node, k1, k2, zrat2, n1 - 1, P);
          const double s1 = zblend(s.salt_face, k1, k2, zrat2, f2 - 1, J);
          const double s3 = zblend(s.salt_face, k1, k2, zrat2, f4 - 1, J);
          tt = 0.0;
          ss = 0.0;
          tt = tt + t0 * sq[0];
          ss = ss + s0 * sq[0];
          tt = tt + t1 * sq[1];
          ss = ss + s1 * sq[1];
          tt = tt + ta * sq[2];
          ss = ss + sa * sq[2];
          tt = tt + t3 * sq[3];
          ss = ss + s3 * sq[3];
        }
      }
    }
  }
  if (!found) return GOMGPU_ERR_ELM_NO_SUBCELL;
  return 0;
}

// One thread per (layer, cell): the backtraced tracers of a wet cell-layer with the surface
// heat source (solve_transport_equation_ELM_v8.f90:40-140) and the explicit horizontal
// diffusion through each face towards the face's second cell (:150-190), into w.tvd_x:
// tracer slot 0 the temperature right-hand side, slot 1 the salt one. As in the reference the
// salt differences enter the temperature row and the temperature differences the salt row.
__global__ void __launch_bounds__(256) k_elm_tracer(DevMesh m, DevState s, DevForcing f, DevWork w, DevDims d,
                                                    double MSL, double dt, double rho_o, int Nt, int *err) {
  const int E = d.E, J = d.J, P = d.P, L = d.L;
  int i, k;
  if (!gom_tile(E, L, k, i)) return;
  k += 1;
  const int tl = s.top_layer_at_element[i], bl = m.bottom_layer_at_element[i];
  if (tl == 0 || k < bl || k > tl) return;
  const int N = m.tri_or_quad[i];
  const double cp = 4186.0;
  double z0;
  if (k == tl)
    z0 = MSL + s.eta_cell[i] - AL(s.dz_cell, k, i, E) * 0.5;
  else
    z0 = m.z_level[k] - AL(s.dz_cell, k, i, E) * 0.5;
  double u_temp = 0.0, v_temp = 0.0;
  int bt_step = 0;
  for (int l = 1; l <= N; l++) {
    const int nd = TB(m.nodenum_at_cell, l, i, E) - 1;
    u_temp = u_temp + (A0(s.u_node, k, nd, P) + A0(s.u_node, k - 1, nd, P)) * 0.5;
    v_temp = v_temp + (A0(s.v_node, k, nd, P) + A0(s.v_node, k - 1, nd, P)) * 0.5;
    bt_step = bt_step + AL(s.num_sub_elm_iteration, k, nd, P);
  }
  u_temp = u_temp / N;
  v_temp = v_temp / N;
  bt_step = bt_step / N;
  double uu = u_temp, vv = v_temp;
  double ww = (A0(s.wn_cell, k, i, E) + A0(s.wn_cell, k - 1, i, E)) * 0.5;
  const double vmag = sqrt(uu * uu + vv * vv + ww * ww);
  double tc, sc;
  if (vmag <= (double)1.e-4f) {
    tc = AL(s.temp_cell, k, i, E);
    sc = AL(s.salt_cell, k, i, E);
  } else {
    const double bt_dt = dt / bt_step;
    ElmCtx c{m, s, d, MSL};
    ElmFoot ft;
    int nnel = i + 1, jlev = k;
    bool near = false;
    int e = backtrace(c, bt_step, bt_dt, uu, vv, ww, m.x_cell[i], m.y_cell[i], z0, nnel, jlev, ft, near);
    if (near) atomicAdd(w.elm_near + 1, 1ull);
    if (!e) e = elm_tracer_foot(m, s, f, d, Nt, nnel, jlev, ft, tc, sc);
    if (e) {
      latch(err, e, i + 1);
      return;
    }
  }
  if (f.phi_sz) tc = tc + dt * AL(f.phi_sz, k, i, E) / (rho_o * cp * AL(s.dz_cell, k, i, E));
  double sum1 = 0.0, sum2 = 0.0;
  const double si = AL(s.salt_cell, k, i, E), ti = AL(s.temp_cell, k, i, E);
  for (int l = 1; l <= N; l++) {
    const int j = TB(m.facenum_at_cell, l, i, E) - 1;
    const int ie = TB(m.adj_cellnum_at_face, 2, j, J);
    if (ie != 0) {
      const double rt = m.face_length[j] * AL(s.dz_face, k, j, J) * AL(s.Kh, k, j, J) / m.delta_j[j];
      sum1 = sum1 + rt * (AL(s.salt_cell, k, ie - 1, E) - si);
      sum2 = sum2 + rt * (AL(s.temp_cell, k, ie - 1, E) - ti);
    }
  }
  sum1 = sum1 * dt / (m.area[i] * AL(s.dz_cell, k, i, E));
  sum2 = sum2 * dt / (m.area[i] * AL(s.dz_cell, k, i, E));
  w.tvd_x[(size_t)(k - 1) * E + i] = tc + sum1;
  w.tvd_x[((size_t)L + (k - 1)) * E + i] = sc + sum2;
}

// Row r (1-based, surface first) of a wet column's vertical diffusion system
// (solve_transport_equation_ELM_v8.f90:200-250); the top row's upper coefficient is
// -dt + Kv/dzhalf as the reference writes it.
__device__ __forceinline__ void elm_row(const DevState &s, int i, int E, int tl, int bl, int r, double dt,
                                        double &av, double &bv, double &cv) {
  const int nv = tl - bl + 1;
  av = 0.0;
  cv = 0.0;
  if (tl == bl) {
    bv = AL(s.dz_cell, tl, i, E);
  } else if (r == 1) {
    cv = -dt + A0(s.Kv, tl - 1, i, E) / A0(s.dzhalf_cell, tl - 1, i, E);
    bv = AL(s.dz_cell, tl, i, E) - cv;
  } else if (r == nv) {
    av = -dt * A0(s.Kv, bl, i, E) / A0(s.dzhalf_cell, bl, i, E);
    bv = -av + AL(s.dz_cell, bl, i, E);
  } else {
    const int k = tl - r + 1;
    av = -dt * A0(s.Kv, k, i, E) / A0(s.dzhalf_cell, k, i, E);
    cv = -dt * A0(s.Kv, k - 1, i, E) / A0(s.dzhalf_cell, k - 1, i, E);
    bv = -av + AL(s.dz_cell, k, i, E) - cv;
  }
}

// One thread per cell column: the Thomas recurrence of tridiagonal_solver.f90 with the
// temperature and salt right-hand sides (dz times the rows of w.tvd_x, solved in place,
// elimination factors in w.tri_gam), the solution into temp/salt_cell_new and the values
// extended below the bottom and above the top layer. A dry column keeps its new values;
// copy_salt / copy_temp say that the new array still holds an older state and stands for
// the old one, which such a column then receives.
__global__ void k_elm_column(DevMesh m, DevState s, DevWork w, DevDims d, double dt, int copy_salt, int copy_temp,
                             int *err) {
  const int i = gom_block() * blockDim.x + threadIdx.x;
  const int E = d.E, L = d.L;
  if (i >= E) return;
  const int tl = s.top_layer_at_element[i], bl = m.bottom_layer_at_element[i];
  if (tl == 0) {
    for (int k = 1; k <= L; k++) {
      if (copy_temp) AL(s.temp_cell_new, k, i, E) = AL(s.temp_cell, k, i, E);
      if (copy_salt) AL(s.salt_cell_new, k, i, E) = AL(s.salt_cell, k, i, E);
    }
    return;
  }
  const int nv = tl - bl + 1;
  double *xt = w.tvd_x, *xs = w.tvd_x + (size_t)L * E;
  double av, bv, cv, cprev;
  elm_row(s, i, E, tl, bl, 1, dt, av, bv, cv);
  if (bv == 0.0) {
    latch(err, GOMGPU_ERR_TRIDIAG_PIVOT, i + 1);
    return;
  }
  double den = bv;
  double dz = AL(s.dz_cell, tl, i, E);
  double vt = dz * AL(xt, tl, i, E) / den, vs = dz * AL(xs, tl, i, E) / den;
  AL(xt, tl, i, E) = vt;
  AL(xs, tl, i, E) = vs;
  cprev = cv;
  for (int r = 2; r <= nv; r++) {
    const int k = tl - r + 1;
    elm_row(s, i, E, tl, bl, r, dt, av, bv, cv);
    const double gam = cprev / den;
    den = bv - av * gam;
    if (den == 0.0) {
      latch(err, GOMGPU_ERR_TRIDIAG_PIVOT, i + 1);
      return;
    }
    dz = AL(s.dz_cell, k, i, E);
    vt = (dz * AL(xt, k, i, E) - av * vt) / den;
    vs = (dz * AL(xs, k, i, E) - av * vs) / den;
    AL(xt, k, i, E) = vt;
    AL(xs, k, i, E) = vs;
    AL(w.tri_gam, r, i, E) = gam;
    cprev = cv;
  }
  // back substitution: layers bl .. tl
  AL(s.temp_cell_new, bl, i, E) = vt;
  AL(s.salt_cell_new, bl, i, E) = vs;
  for (int k = 1; k < bl; k++) {
    AL(s.temp_cell_new, k, i, E) = vt;
    AL(s.salt_cell_new, k, i, E) = vs;
  }
  for (int r = nv - 1; r >= 1; r--) {
    const int k = tl - r + 1;
    const double gam = AL(w.tri_gam, r + 1, i, E);
    vt = AL(xt, k, i, E) - gam * vt;
    vs = AL(xs, k, i, E) - gam * vs;
    AL(s.temp_cell_new, k, i, E) = vt;
    AL(s.salt_cell_new, k, i, E) = vs;
  }
  for (int k = tl + 1; k <= L; k++) {
    AL(s.temp_cell_new, k, i, E) = vt;
    AL(s.salt_cell_new, k, i, E) = vs;
  }
}

// Node values (solve_transport_equation_ELM_v8.f90:270-320), one thread per node, top down:
// the area*dz_new weighted mean of the adjacent cells wet at the level; a level whose
// adjacent cells are all dry takes the value of the level above (unless it is the node's
// top level); then the values extended below the bottom and above the top. Dry nodes keep
// their values.
__global__ void k_elm_node(DevMesh m, DevState s, DevDims d) {
  const int i = gom_block() * blockDim.x + threadIdx.x;
  const int P = d.P, E = d.E, L = d.L;
  if (i >= P) return;
  const int tn = s.top_layer_at_node[i], bn = m.bottom_layer_at_node[i];
  if (tn == 0) return;
  const int nadj = m.adj_cells_at_node[i];
  double tabove = 0.0, sabove = 0.0;
  for (int k = tn; k >= bn; k--) {
    double sum0 = 0.0, sum1 = 0.0, sum2 = 0.0;
    int icount = 0;
    for (int l = 1; l <= nadj; l++) {
      const int ie = TB(m.adj_cellnum_at_node, l, i, P) - 1;
      if (k >= m.bottom_layer_at_element[ie] && k <= s.top_layer_at_element[ie]) {
        const double wgt = m.area[ie] * AL(s.dz_cell_new, k, ie, E);
        sum0 = sum0 + wgt;
        sum1 = sum1 + wgt * AL(s.temp_cell_new, k, ie, E);
        sum2 = sum2 + wgt * AL(s.salt_cell_new, k, ie, E);
      } else {
        icount++;
      }
    }
    double tv = sum0 == 0.0 ? 0.0 : sum1 / sum0, sv = sum0 == 0.0 ? 0.0 : sum2 / sum0;
    if (icount == nadj && k != tn) {
      tv = tabove;
      sv = sabove;
    }
    AL(s.temp_node, k, i, P) = tv;
    AL(s.salt_node, k, i, P) = sv;
    tabove = tv;
    sabove = sv;
  }
  for (int k = 1; k < bn; k++) {
    AL(s.temp_node, k, i, P) = tabove;
    AL(s.salt_node, k, i, P) = sabove;
  }
  const double ttop = AL(s.temp_node, tn, i, P), stop = AL(s.salt_node, tn, i, P);
  for (int k = tn + 1; k <= L; k++) {
    AL(s.temp_node, k, i, P) = ttop;
    AL(s.salt_node, k, i, P) = stop;
  }
}

// Face values: the mean of the face's two nodes at every layer (:325-335).
__global__ void k_elm_face(DevMesh m, DevState s, DevDims d) {
  const long q = (long)gom_block() * blockDim.x + threadIdx.x;
  const int J = d.J, P = d.P;
  if (q >= (long)J * d.L) return;
  const int j = (int)(q % J), k = (int)(q / J) + 1;
  const int n1 = TB(m.nodenum_at_face, 1, j, J) - 1, n2 = TB(m.nodenum_at_face, 2, j, J) - 1;
  AL(s.salt_face, k, j, J) = (AL(s.salt_node, k, n1, P) + AL(s.salt_node, k, n2, P)) * 0.5;
  AL(s.temp_face, k, j, J) = (AL(s.temp_node, k, n1, P) + AL(s.temp_node, k, n2, P)) * 0.5;
}

// ================================================================ momentum
// solve_momentum_equation.f90:67-655, in two passes. k_mom_rhs: one thread per (row, face)
// builds the right-hand sides G1, G2 of tridiagonal row r (local index, surface first) and
// Kh; k_mom_solve: one thread per face runs the Thomas recurrence of
// tridiagonal_solver.f90 with the three right-hand sides [G1, G2, dz], in place in the
// output arrays (the elimination factors go to w.tri_gam). Every value is computed by the
// same expression sequence as the reference's per-face loop.
// Cell-layer means of the node velocities at mid-layer (the cell terms of the Smagorinsky
// stencil, solve_momentum_equation.f90:280-300), in the reference's node order; one thread
// per (layer, cell).
__global__ void k_cell_uv(DevMesh m, DevState s, DevWork w, DevDims d) {
  // GOM_CUV_LAYERS consecutive layers of one cell per thread: the node ids are read once and
  // each node level serves the two layers it bounds
  const int E = d.E, P = d.P, L = d.L;
  const int nch = (L + GOM_CUV_LAYERS - 1) / GOM_CUV_LAYERS;
  long t = (long)gom_block() * blockDim.x + threadIdx.x;
  if (t >= (long)nch * E) return;
  const int ie = (int)(t % E), k1 = (int)(t / E) * GOM_CUV_LAYERS + 1;
  const int N = m.tri_or_quad[ie];
  int nd[4];
#pragma unroll
  for (int ll = 0; ll < 4; ll++) nd[ll] = ll < N ? TB(m.nodenum_at_cell, ll + 1, ie, E) - 1 : nd[0];
  double ul[4], vl[4];  // the nodes' values at the level below the current layer
#pragma unroll
  for (int ll = 0; ll < 4; ll++) {
    ul[ll] = A0(s.u_node, k1 - 1, nd[ll], P);
    vl[ll] = A0(s.v_node, k1 - 1, nd[ll], P);
  }
#pragma unroll
  for (int z = 0; z < GOM_CUV_LAYERS; z++) {
    const int kk = k1 + z;
    if (kk > L) continue;
    double ut = 0.0, vt = 0.0;
#pragma unroll
    for (int ll = 0; ll < 4; ll++) {
      const double uu = A0(s.u_node, kk, nd[ll], P), vv = A0(s.v_node, kk, nd[ll], P);
      if (ll < N) {
        ut = ut + (uu + ul[ll]) * 0.5 / N;
        vt = vt + (vv + vl[ll]) * 0.5 / N;
      }
      ul[ll] = uu;
      vl[ll] = vv;
    }
    AL(w.cell_u, kk, ie, E) = ut;
    AL(w.cell_v, kk, ie, E) = vt;
  }
}

// One block covers 64 consecutive faces (lane = face) and all their rows: wave v of the block
// handles rows v+1, v+1+W, ... (W waves). The baroclinic sums (solve_momentum_equation.f90:
// 400-460) of row k add, for the layers k2 = kk..top, (rho2 - rho1) * dz_face(k2) -- halved
// at k2 = kk -- where the layer is wet on both sides. Those products depend on (face, k2)
// only, so the block first forms them once per layer into LDS (zero where a side is dry:
// adding +0.0 to a sum that starts at +0.0 changes nothing, and halving the product equals
// halving the density difference first, both exact), and each row then adds them from LDS in
// the reference's order: one gather per layer instead of one per (row, layer).
__global__ void __launch_bounds__(256) GOM_WAVES(GOM_WAVES_MOM_RHS) k_mom_rhs(DevMesh m, DevState s, DevForcing f, DevWork w, DevDims d,
                                                 MomParams q) {
  extern __shared__ double bct[];  // [2][L][64] baroclinic products (cells, nodes)
  const int J = d.J, E = d.E, P = d.P, L = d.L;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int j = gom_block() * 64 + lane;
  const bool valid = j < J;
  const int kt = valid ? s.top_layer_at_face[j] : 0;
  const int kb = valid ? m.bottom_layer_at_face[j] : 1;
  const int node1 = valid ? TB(m.nodenum_at_face, 1, j, J) - 1 : 0;
  const int node2 = valid ? TB(m.nodenum_at_face, 2, j, J) - 1 : 0;
  const int a1 = valid ? TB(m.adj_cellnum_at_face, 1, j, J) : 1;
  const int a2 = valid ? TB(m.adj_cellnum_at_face, 2, j, J) : 0;
  if (q.baroclinic == 1) {
    int cb1 = 0, ct1 = -1, cb2 = 0, ct2c = -1, nb1 = 0, nt1 = -1, nb2 = 0, nt2 = -1;
    if (valid && kt != 0) {
      if (a2 != 0) {
        cb1 = m.bottom_layer_at_element[a1 - 1];
        ct1 = s.top_layer_at_element[a1 - 1];
        cb2 = m.bottom_layer_at_element[a2 - 1];
        ct2c = s.top_layer_at_element[a2 - 1];
      }
      nb1 = m.bottom_layer_at_node[node1];
      nt1 = s.top_layer_at_node[node1];
      nb2 = m.bottom_layer_at_node[node2];
      nt2 = s.top_layer_at_node[node2];
    }
    for (int k2 = wv + 1; k2 <= L; k2 += nw) {
      double t1 = 0.0, t2 = 0.0;
      if (valid && kt != 0 && k2 >= kb && k2 <= kt) {
        double dz2 = AL(s.dz_face, k2, j, J);
        if (a2 != 0 && k2 >= cb1 && k2 <= ct1 && k2 >= cb2 && k2 <= ct2c)
          t1 = (AL(s.rho_cell, k2, a2 - 1, E) - AL(s.rho_cell, k2, a1 - 1, E)) * dz2;
        if (k2 >= nb1 && k2 <= nt1 && k2 >= nb2 && k2 <= nt2)
          t2 = (AL(s.rho_node, k2, node2, P) - AL(s.rho_node, k2, node1, P)) * dz2;
      }
      bct[(k2 - 1) * 64 + lane] = t1;
      bct[(L + k2 - 1) * 64 + lane] = t2;
    }
    __syncthreads();
  }
  if (!valid || kt == 0) return;
  const int n = kt - kb + 1;
  const double dt = q.dt, g = q.g, th = q.th;
  const int bt = m.boundary_type_of_face[j];
  const double dj = m.delta_j[j], Lf = m.face_length[j];
  const double ct = m.cos_theta[j], st = m.sin_theta[j], ct2 = m.cos_theta2[j], st2 = m.sin_theta2[j];
  const double cor = m.coriolis_factor[j];
  // air pressure differences are layer independent
  double airc = 0.0;
  if (a2 != 0) {
    double tp[2];
    for (int l = 0; l < 2; l++) {
      int ie = TB(m.adj_cellnum_at_face, l + 1, j, J) - 1;
      int N = m.tri_or_quad[ie];
      tp[l] = 0.0;
      for (int ll = 1; ll <= N; ll++) tp[l] = tp[l] + f.air_p1[TB(m.nodenum_at_cell, ll, ie, E) - 1] / N;
    }
    airc = -(dt / q.rho_o * (tp[1] - tp[0]) / dj);
  }
  const double rtemp3a = -(dt / q.rho_o * ((f.air_p1[node2] - f.air_p1[node1]) / Lf));
  const bool wet_nodes = (s.wetdry_node[node1] == 0 && s.wetdry_node[node2] == 0);
  const double spinup_bc = f.scal[0];
  const int ibnd = m.isflowside3[j];
  for (int k = wv + 1; k <= n; k += nw) {
    double rtemp0 = 0.0, rtemp1, rtemp2 = 0.0, rtemp3, rtemp4;
    int kk = kt - k + 1;
    double dz = AL(s.dz_face, kk, j, J);
    double unf = AL(s.un_face, kk, j, J), vnf = AL(s.vn_face, kk, j, J);
    double G1 = dz * (AL(s.un_ELM, kk, j, J) + cor * dt * vnf);
    double G2 = dz * (AL(s.vn_ELM, kk, j, J) - cor * dt * unf);
    if (a2 != 0) {
      double temp[13];
      for (int l = 1; l <= 2; l++) {
        int nd = TB(m.nodenum_at_face, l, j, J) - 1;
        double ut = (A0(s.u_node, kk, nd, P) + A0(s.u_node, kk - 1, nd, P)) * 0.5;
        double vt = (A0(s.v_node, kk, nd, P) + A0(s.v_node, kk - 1, nd, P)) * 0.5;
        temp[2 * l - 1] = ut * ct + vt * st;
        temp[2 * l] = -ut * st + vt * ct;
      }
      for (int l = 1; l <= 2; l++) {
        int ie = TB(m.adj_cellnum_at_face, l, j, J) - 1;
        double ut = AL(w.cell_u, kk, ie, E), vt = AL(w.cell_v, kk, ie, E);
        temp[2 * l + 3] = ut * ct + vt * st;
        temp[2 * l + 4] = -ut * st + vt * ct;
      }
      double dudx = (temp[7] - temp[5]) / dj, dvdx = (temp[8] - temp[6]) / dj;
      double dudy = (temp[3] - temp[1]) / Lf, dvdy = (temp[4] - temp[2]) / Lf;
      double sx = dvdx + dudy;
      double hdiff = q.smag * sqrt(dudx * dudx + dvdy * dvdy + 0.5 * (sx * sx));
      double kh = hdiff * m.area[a1 - 1];
      hdiff = hdiff + q.Ah0;
      AL(s.Kh, k, j, J) = kh + q.Kh0;
      double d2udy = 4.0 * (temp[3] + temp[1] - 2 * unf) / (Lf * Lf);
      double d2vdy = 4.0 * (temp[4] + temp[2] - 2 * vnf) / (Lf * Lf);
      for (int l = 1; l <= 2; l++) {
        int ie = TB(m.adj_cellnum_at_face, l, j, J) - 1;
        double ddx = m.x_cell[ie] - m.x_face[j], ddy = m.y_cell[ie] - m.y_face[j];
        double rl = sqrt(ddx * ddx + ddy * ddy);
        temp[2 * l + 7] = (temp[2 * l + 3] - unf) / rl;
        temp[2 * l + 8] = (temp[2 * l + 4] - vnf) / rl;
      }
      double d2udx = 2.0 * (temp[9] + temp[11]) / dj;
      double d2vdx = 2.0 * (temp[10] + temp[12]) / dj;
      G1 = G1 + dz * dt * hdiff * (d2udx + d2udy);
      G2 = G2 + dz * dt * hdiff * (d2vdx + d2vdy);
    }
    if (bt == 0) {
      rtemp0 = g * dt / dj * (1.0 - th) * (s.eta_cell[a2 - 1] - s.eta_cell[a1 - 1]);
      rtemp1 = rtemp0 * ct2;
      rtemp2 = rtemp0 * st2;
      G1 = G1 - dz * rtemp1;
    } else if (bt > 0) {
      int ie = m.ob_element_flag[a1 - 1];
      rtemp1 = g * dt / dj * (1.0 - th) * (f.eta_at_ob_old[ie - 1] - s.eta_cell[a1 - 1]);
      G1 = G1 - dz * rtemp1;
    } else if (bt == -1) {
      G1 = 0.0;
    }
    if (wet_nodes) {
      rtemp0 = g * dt / Lf * (1.0 - th) * (s.eta_node[node2] - s.eta_node[node1]);
      if (bt == 0)
        G2 = G2 - dz * (rtemp0 + rtemp2);
      else
        G2 = G2 - dz * rtemp0;
    } else if (a2 == 0) {
      rtemp0 = 0.0;  // defined value for the reference's stale private variable
    }
    if (q.baroclinic == 1) {
      double sum1 = 0.0, sum2 = 0.0;
      for (int k2 = kk; k2 <= kt; k2++) {
        double fct = (k2 == kk) ? 0.5 : 1.0;
        sum1 = sum1 + fct * bct[(k2 - 1) * 64 + lane];
        sum2 = sum2 + fct * bct[(L + k2 - 1) * 64 + lane];
      }
      double dzl = AL(s.dz_face, k, j, J);
      rtemp1 = dzl * (-(g * dt / (q.rho_o * dj))) * sum1;
      rtemp2 = dzl * (-(g * dt / (q.rho_o * Lf))) * sum2;
      rtemp3 = rtemp1 * ct2;
      rtemp4 = rtemp1 * st2;
      G1 = G1 + rtemp3 * spinup_bc;
      G2 = G2 + (rtemp2 + rtemp4) * spinup_bc;
    }
    if (k == 1) {
      G1 = G1 + dt * f.wind_stress_normal[j];
      G2 = G2 + dt * f.wind_stress_tangnt[j];
    }
    G1 = G1 + 0.0;  // equilibrium tide term (zero without tidal potential species)
    G2 = G2 + 0.0;
    if (a2 != 0) rtemp0 = airc;
    rtemp1 = rtemp0 * ct2;
    rtemp2 = rtemp0 * st2;
    G1 = G1 + dz * rtemp1;
    rtemp3 = rtemp3a;
    G2 = G2 + dz * (rtemp3 + rtemp2);
    if (bt == -1) G1 = 0.0;
    if (ibnd > 0) G1 = G1 + dz * f.Qu_boundary[ibnd - 1];
    AL(s.AinvG1, k, j, J) = G1;
    AL(s.AinvG2, k, j, J) = G2;
  }
}

// Tridiagonal row r (0-based, surface first) of face j: solve_momentum_equation.f90:120-169.
__device__ __forceinline__ void mom_row(const DevState &s, const double *gamma_t, int j, int J, int kt, int kb,
                                        int r, double dt, double &a, double &b, double &c) {
  int n = kt - kb + 1;
  a = 0.0;
  c = 0.0;
  if (kt == kb) {
    b = AL(s.dz_face, kt, j, J) + s.Gamma_B[j] * dt;
  } else if (r == 0) {
    c = -(dt * A0(s.Av, kt - 1, j, J) / A0(s.dzhalf_face, kt - 1, j, J));
    b = gamma_t[j] * dt + AL(s.dz_face, kt, j, J) - c;
  } else if (r == n - 1) {
    a = -(dt * A0(s.Av, kb, j, J) / A0(s.dzhalf_face, kb, j, J));
    b = -a + AL(s.dz_face, kb, j, J) + s.Gamma_B[j] * dt;
  } else {
    int k = kt - r;
    a = -(dt * A0(s.Av, k, j, J) / A0(s.dzhalf_face, k, j, J));
    c = -(dt * A0(s.Av, k - 1, j, J) / A0(s.dzhalf_face, k - 1, j, J));
    b = -a + AL(s.dz_face, k, j, J) - c;
  }
}

#define MS_CHUNK 4
__global__ void GOM_WAVES(GOM_WAVES_MOM_SOLVE) k_mom_solve(DevMesh m, DevState s, DevForcing f, DevWork w, DevDims d, MomParams q, int *err) {
  int j = gom_block() * blockDim.x + threadIdx.x;
  const int J = d.J, L = d.L;
  if (j >= J) return;
  int kt = s.top_layer_at_face[j];
  if (kt == 0) {
    for (int k = 1; k <= L; k++) {
      AL(s.AinvG1, k, j, J) = 0.0;
      AL(s.AinvG2, k, j, J) = 0.0;
      AL(s.AinvDeltaZ1, k, j, J) = 0.0;
      AL(s.AinvDeltaZ2, k, j, J) = 0.0;
    }
    return;
  }
  const double dt = q.dt;
  int kb = m.bottom_layer_at_face[j];
  int n = kt - kb + 1;
  int bt = m.boundary_type_of_face[j];
  // forward elimination (rows 0..n-1, RHS 3 is dz of the row's layer). Row r >= 1 (layer
  // k = kt - r) has a = the c of row r-1 (the same expression of the same inputs), c from
  // layer k-1, b from its own dz. Rows are taken in chunks of MS_CHUNK whose inputs are all
  // loaded (branch-free, indexes clamped into the column) before the recurrence runs, so a
  // thread waits for one memory latency per chunk rather than per row.
  double a, b, c, cprev;
  mom_row(s, f.Gamma_T, j, J, kt, kb, 0, dt, a, b, c);
  int e = 0;
  double den = b;
  double x0, x1, x2;
  if (den == 0.0) {
    e = GOMGPU_ERR_TRIDIAG_PIVOT;
  } else {
    x0 = AL(s.AinvG1, 1, j, J) / den;
    x1 = AL(s.AinvG2, 1, j, J) / den;
    x2 = AL(s.dz_face, kt, j, J) / den;
    AL(s.AinvG1, 1, j, J) = x0;
    AL(s.AinvG2, 1, j, J) = x1;
    AL(s.AinvDeltaZ2, 1, j, J) = x2;
    cprev = c;
    const double gbdt = s.Gamma_B[j] * dt;
    for (int r0 = 1; r0 < n && !e; r0 += MS_CHUNK) {
      double avm[MS_CHUNK], dzhm[MS_CHUNK], dzk[MS_CHUNK], g1[MS_CHUNK], g2[MS_CHUNK];
#pragma unroll
      for (int i = 0; i < MS_CHUNK; i++) {
        const int r = r0 + i < n ? r0 + i : n - 1, k = kt - r;
        avm[i] = A0(s.Av, k - 1, j, J);
        dzhm[i] = A0(s.dzhalf_face, k - 1, j, J);
        dzk[i] = AL(s.dz_face, k, j, J);
        g1[i] = AL(s.AinvG1, r + 1, j, J);
        g2[i] = AL(s.AinvG2, r + 1, j, J);
      }
#pragma unroll
      for (int i = 0; i < MS_CHUNK; i++) {
        const int r = r0 + i;
        if (r >= n || e) continue;
        a = cprev;
        c = r < n - 1 ? -(dt * avm[i] / dzhm[i]) : 0.0;
        b = r == n - 1 ? -a + dzk[i] + gbdt : -a + dzk[i] - c;
        double gam = cprev / den;
        den = b - a * gam;
        if (den == 0.0) {
          e = GOMGPU_ERR_TRIDIAG_PIVOT;
          continue;
        }
        x0 = (g1[i] - a * x0) / den;
        x1 = (g2[i] - a * x1) / den;
        x2 = (dzk[i] - a * x2) / den;
        AL(s.AinvG1, r + 1, j, J) = x0;
        AL(s.AinvG2, r + 1, j, J) = x1;
        AL(s.AinvDeltaZ2, r + 1, j, J) = x2;
        AL(w.tri_gam, r + 1, j, J) = gam;
        cprev = c;
      }
    }
  }
  if (e) {
    latch(err, e, j + 1);
    return;
  }
  // back substitution, in chunks likewise
  for (int r0 = n - 2; r0 >= 0; r0 -= MS_CHUNK) {
    double gm[MS_CHUNK], h0[MS_CHUNK], h1[MS_CHUNK], h2[MS_CHUNK];
#pragma unroll
    for (int i = 0; i < MS_CHUNK; i++) {
      const int r = r0 - i >= 0 ? r0 - i : 0;
      gm[i] = AL(w.tri_gam, r + 2, j, J);
      h0[i] = AL(s.AinvG1, r + 1, j, J);
      h1[i] = AL(s.AinvG2, r + 1, j, J);
      h2[i] = AL(s.AinvDeltaZ2, r + 1, j, J);
    }
#pragma unroll
    for (int i = 0; i < MS_CHUNK; i++) {
      const int r = r0 - i;
      if (r < 0) continue;
      x0 = h0[i] - gm[i] * x0;
      x1 = h1[i] - gm[i] * x1;
      x2 = h2[i] - gm[i] * x2;
      AL(s.AinvG1, r + 1, j, J) = x0;
      AL(s.AinvG2, r + 1, j, J) = x1;
      AL(s.AinvDeltaZ2, r + 1, j, J) = x2;
    }
  }
  bool land = (bt == -1);
  bool zero_g1 = land && m.isflowside3[j] == 0;
  for (int k = 1; k <= L; k++) {
    if (k <= n) {
      if (zero_g1) AL(s.AinvG1, k, j, J) = 0.0;
      double z = AL(s.AinvDeltaZ2, k, j, J);
      AL(s.AinvDeltaZ1, k, j, J) = land ? 0.0 : z;
    } else {
      AL(s.AinvG1, k, j, J) = 0.0;
      AL(s.AinvG2, k, j, J) = 0.0;
      AL(s.AinvDeltaZ1, k, j, J) = 0.0;
      AL(s.AinvDeltaZ2, k, j, J) = 0.0;
    }
  }
}

// ================================================================ free surface
// Per-face column sums of solve_free_surface_equation.f90:79-86, 140-146 (top down).
__global__ void k_fs_face(DevMesh m, DevState s, DevWork w, DevDims d, double gdt2th2) {
  int j = gom_block() * blockDim.x + threadIdx.x;
  const int J = d.J;
  if (j >= J) return;
  int kt = s.top_layer_at_face[j];
  double sdz = 0.0, sun = 0.0, sg = 0.0;
  if (kt != 0) {
    int n = kt - m.bottom_layer_at_face[j] + 1;
    for (int k = 1; k <= n; k++) {
      int kk = kt + 1 - k;
      double dz = AL(s.dz_face, kk, j, J);
      sdz = sdz + dz * AL(s.AinvDeltaZ1, k, j, J);
      sun = sun + dz * AL(s.un_face, kk, j, J);
      sg = sg + dz * AL(s.AinvG1, k, j, J);
    }
    sdz = gdt2th2 * m.face_length[j] / m.delta_j[j] * sdz;
  }
  w.face_sdz[j] = sdz;
  w.face_sun[j] = sun;
  w.face_sg[j] = sg;
}

// Matrix row and right-hand side per cell: solve_free_surface_equation.f90:65-214.
// The row is kept as ELL (diagonal + one slot per cell face, neighbour order of the
// reference CSR); the initial guess is eta_cell_new (:241).
__global__ void k_fs_cell(DevMesh m, DevState s, DevForcing f, DevWork w, DevDims d, double dt,
                          double th, int *err) {
  int i = gom_block() * blockDim.x + threadIdx.x;
  const int E = d.E;
  if (i >= E) return;
  int N = m.tri_or_quad[i];
  double diag = m.area[i];
  double off[4] = {0.0, 0.0, 0.0, 0.0};
  for (int l = 1; l <= N; l++) {
    int j = TB(m.facenum_at_cell, l, i, E) - 1;
    if (s.top_layer_at_face[j] != 0) {
      double c = w.face_sdz[j] * m.cos_theta2[j];
      if (c < 0.0) latch(err, GOMGPU_ERR_NOT_SPD, i + 1);
      diag = diag + c;
      if (TB(m.adj_cellnum_at_cell, l, i, E) != 0) off[l - 1] = -c;
    }
  }
  double rhs = m.area[i] * s.eta_cell[i];
  for (int l = 1; l <= N; l++) {
    int j = TB(m.facenum_at_cell, l, i, E) - 1;
    if (s.top_layer_at_face[j] != 0) {
      double sg = (double)TB(m.sign_in_outflow, l, i, E), Lf = m.face_length[j];
      rhs = rhs - (1.0 - th) * dt * sg * Lf * w.face_sun[j] - th * dt * sg * Lf * w.face_sg[j];
    }
  }
  int obf = m.ob_element_flag[i];
  if (obf > 0) {
    for (int l = 1; l <= N; l++) {
      int j = TB(m.facenum_at_cell, l, i, E) - 1;
      if (m.boundary_type_of_face[j] > 0 && s.top_layer_at_face[j] != 0)
        rhs = rhs + w.face_sdz[j] * f.eta_at_ob_new[obf - 1];
    }
  }
  w.fs_diag[i] = diag;
  for (int l = 0; l < 4; l++) w.fs_off[ix(l, i, E)] = off[l];
  w.fs_rhs[i] = rhs;
  w.pcg_x[i] = s.eta_cell_new[i];
}

// Area-weighted node elevation and wet/dry flag: solve_free_surface_equation.f90:353-374
__global__ void k_eta_node(DevMesh m, DevState s, DevDims d) {
  int i = gom_block() * blockDim.x + threadIdx.x;
  const int P = d.P;
  if (i >= P) return;
  int wet = 0;
  double sum1 = 0.0, sum0 = 0.0;
  int nadj = m.adj_cells_at_node[i];
  for (int jj = 1; jj <= nadj; jj++) {
    int ie = TB(m.adj_cellnum_at_node, jj, i, P) - 1;
    if (s.top_layer_at_element[ie] == 0) {
      wet = 1;
    } else {
      sum1 = sum1 + m.area[ie] * s.eta_cell_new[ie];
      sum0 = sum0 + m.area[ie];
    }
  }
  s.wetdry_node[i] = wet;
  if (wet == 0) s.eta_node[i] = sum1 / sum0;
}

// Dry-node fill from wet neighbours: solve_free_surface_equation.f90:378-394
__global__ void k_eta_node_dry(DevMesh m, DevState s, DevDims d) {
  int i = gom_block() * blockDim.x + threadIdx.x;
  const int P = d.P;
  if (i >= P) return;
  if (s.wetdry_node[i] != 1) return;
  double sum0 = 0.0;
  int icount = 0;
  int na = m.adj_nodes_at_node[i];
  for (int jj = 1; jj <= na; jj++) {
    int nd = m.adj_nodenum_at_node[ix(jj, i, P)] - 1;  // table lower bound 0
    if (s.wetdry_node[nd] == 0) {
      icount++;
      sum0 = sum0 + s.eta_node[nd];
    }
  }
  if (icount != 0) s.eta_node[i] = sum0 / icount;
}

// ================================================================ velocities
// solve_velocities.f90:23-185; one thread per face column. Entries the reference leaves
// untouched (vn_face_new where an end node is dry) keep their previous value, which after
// update_variables equals vn_face; they are written back from vn_face so the arrays can be
// double-buffered.
__global__ void k_velocity_face(DevMesh m, DevState s, DevForcing f, DevDims d, double thgdt) {
  int j = gom_block() * blockDim.x + threadIdx.x;
  const int J = d.J, L = d.L;
  if (j >= J) return;
  int kt = s.top_layer_at_face[j];
  if (kt == 0) {
    for (int k = 1; k <= L; k++) {
      AL(s.un_face_new, k, j, J) = 0.0;
      AL(s.vn_face_new, k, j, J) = 0.0;
    }
    return;
  }
  int n1 = TB(m.nodenum_at_face, 1, j, J) - 1, n2 = TB(m.nodenum_at_face, 2, j, J) - 1;
  int kb = m.bottom_layer_at_face[j];
  int bt = m.boundary_type_of_face[j];
  int a1 = TB(m.adj_cellnum_at_face, 1, j, J), a2 = TB(m.adj_cellnum_at_face, 2, j, J);
  bool wet = (s.wetdry_node[n1] == 0 && s.wetdry_node[n2] == 0);
  bool clampv = (m.initial_wetdry_node[n1] == 1 || m.initial_wetdry_node[n2] == 1);
  double dj = m.delta_j[j], Lf = m.face_length[j];
  double deta = 0.0;
  if (bt == 0) deta = s.eta_cell_new[a2 - 1] - s.eta_cell_new[a1 - 1];
  else if (bt > 0) deta = f.eta_at_ob_new[m.ob_element_flag[a1 - 1] - 1] - s.eta_cell_new[a1 - 1];
  double detan = s.eta_node[n2] - s.eta_node[n1];
  int ibnd = m.isflowside3[j];
  double qriv = ibnd > 0 ? f.Qu_boundary[ibnd - 1] : 0.0;
  double utop = 0.0, vtop = 0.0;
  for (int k = 1; k <= L; k++) {
    double un, vn;
    if (k < kb) {
      un = 0.0;
      vn = 0.0;
    } else if (k <= kt) {
      int kl = kt + 1 - k;  // local (top-down) index
      double rtemp2 = 0.0;
      if (bt == -1) {
        un = 0.0;
      } else if (bt == 0) {
        double r0 = thgdt / dj * deta * AL(s.AinvDeltaZ1, kl, j, J);
        double r1 = r0 * m.cos_theta2[j];
        rtemp2 = r0 * m.sin_theta2[j];
        un = AL(s.AinvG1, kl, j, J) - r1;
      } else {
        double r0 = thgdt / dj * deta * AL(s.AinvDeltaZ1, kl, j, J);
        un = AL(s.AinvG1, kl, j, J) - r0;
      }
      if (wet) {
        double r1 = thgdt / Lf * detan * AL(s.AinvDeltaZ2, kl, j, J);
        vn = (bt == 0) ? AL(s.AinvG2, kl, j, J) - (r1 + rtemp2) : AL(s.AinvG2, kl, j, J) - r1;
      } else {
        vn = AL(s.vn_face, k, j, J);
      }
      if (clampv) vn = dmaxf(-5.0, dminf(vn, 5.0));
      if (ibnd > 0) un = qriv;
      if (k == kt) {
        utop = un;
        vtop = vn;
      }
    } else {
      un = utop;
      vn = vtop;
    }
    AL(s.un_face_new, k, j, J) = un;
    AL(s.vn_face_new, k, j, J) = vn;
  }
}

// Vertical velocity per cell column: solve_velocities.f90:190-258
__global__ void k_velocity_w(DevMesh m, DevState s, DevDims d) {
  int i = gom_block() * blockDim.x + threadIdx.x;
  const int E = d.E, J = d.J, L = d.L;
  if (i >= E) return;
  int t = s.top_layer_at_element[i], b = m.bottom_layer_at_element[i];
  if (t == 0) {
    for (int k = 0; k <= L; k++) A0(s.wn_cell_new, k, i, E) = 0.0;
    return;
  }
  int N = m.tri_or_quad[i];
  int fj[4], fb[4], ft[4];
  double fs[4];
  for (int l = 0; l < N; l++) {
    fj[l] = TB(m.facenum_at_cell, l + 1, i, E) - 1;
    fb[l] = m.bottom_layer_at_face[fj[l]];
    ft[l] = s.top_layer_at_face[fj[l]];
    fs[l] = (double)TB(m.sign_in_outflow, l + 1, i, E) * m.face_length[fj[l]];
  }
  double ar = m.area[i];
  for (int k = 0; k <= b - 2; k++) A0(s.wn_cell_new, k, i, E) = 0.0;
  double w = 0.0;
  A0(s.wn_cell_new, b - 1, i, E) = 0.0;
  for (int k = b; k <= t; k++) {
    double sum1 = 0.0;
    for (int l = 0; l < N; l++)
      if (k >= fb[l] && k <= ft[l]) sum1 = sum1 + fs[l] * AL(s.dz_face, k, fj[l], J) * AL(s.un_face_new, k, fj[l], J);
    w = w - sum1 / ar;
    A0(s.wn_cell_new, k, i, E) = w;
  }
  for (int k = t + 1; k <= L; k++) A0(s.wn_cell_new, k, i, E) = w;
}

// ================================================================ layers
// update_vertical_layers_1.f90:40-133 (cells)
__global__ void k_layers_cell(DevMesh m, DevState s, DevDims d, double MSL, double dry, int *err) {
  int i = gom_block() * blockDim.x + threadIdx.x;
  const int E = d.E, L = d.L;
  if (i >= E) return;
  int tnew = 0;
  double eta = s.eta_cell_new[i];
  double tw = m.h_cell[i] + eta;
  int bl = m.bottom_layer_at_element[i];
  if (tw > dry) {
    double tmp = MSL + eta;
    for (int k = L - 1; k >= 0; k--) {
      if (tmp > m.z_level[k] && tmp <= m.z_level[k + 1]) {
        tnew = k + 1;
        AL(s.dz_cell_new, tnew, i, E) = tmp - m.z_level[k];
        break;
      }
    }
    if (tnew < bl) {
      latch(err, GOMGPU_ERR_LAYER_CELL, i + 1);
      return;
    }
    if (tnew == bl) {
      AL(s.dz_cell_new, tnew, i, E) = tw;
      A0(s.dzhalf_cell_new, tnew, i, E) = tw;
    } else {
      AL(s.dz_cell_new, bl, i, E) = m.z_level[bl] - (MSL - m.h_cell[i]);
      for (int k = bl + 1; k <= tnew - 1; k++) AL(s.dz_cell_new, k, i, E) = m.delta_z[k - 1];
      for (int k = bl; k <= tnew - 1; k++)
        A0(s.dzhalf_cell_new, k, i, E) = (AL(s.dz_cell_new, k + 1, i, E) + AL(s.dz_cell_new, k, i, E)) * 0.5;
      A0(s.dzhalf_cell_new, tnew, i, E) = AL(s.dz_cell_new, tnew, i, E);
      A0(s.dzhalf_cell_new, bl - 1, i, E) = AL(s.dz_cell_new, bl, i, E);
    }
  }
  if (s.top_layer_at_element[i] == 0 && tnew != 0)
    for (int k = 0; k <= L; k++) A0(s.wn_cell_new, k, i, E) = 0.0;
  s.top_layer_at_element[i] = tnew;
}

// update_vertical_layers_1.f90:138-198 (faces)
__global__ void k_layers_face(DevMesh m, DevState s, DevDims d, double MSL, double dry, int *err) {
  int j = gom_block() * blockDim.x + threadIdx.x;
  const int J = d.J, L = d.L;
  if (j >= J) return;
  int tnew = 0;
  int a1 = TB(m.adj_cellnum_at_face, 1, j, J), a2 = TB(m.adj_cellnum_at_face, 2, j, J);
  double eta_max = (a2 != 0) ? dmaxf(s.eta_cell_new[a1 - 1], s.eta_cell_new[a2 - 1]) : s.eta_cell_new[a1 - 1];
  double tw = m.h_face[j] + eta_max;
  int bl = m.bottom_layer_at_face[j];
  if (tw > dry) {
    double tmp = MSL + eta_max;
    for (int k = L - 1; k >= 0; k--) {
      if (tmp > m.z_level[k] && tmp <= m.z_level[k + 1]) {
        tnew = k + 1;
        AL(s.dz_face_new, tnew, j, J) = tmp - m.z_level[k];
        break;
      }
    }
    if (tnew < bl) {
      latch(err, GOMGPU_ERR_LAYER_FACE, j + 1);
      return;
    }
    if (tnew == bl) {
      AL(s.dz_face_new, tnew, j, J) = tw;
      A0(s.dzhalf_face_new, tnew - 1, j, J) = tw;
      A0(s.dzhalf_face_new, tnew, j, J) = tw;
    } else {
      AL(s.dz_face_new, bl, j, J) = m.z_level[bl] - (MSL - m.h_face[j]);
      for (int k = bl + 1; k <= tnew - 1; k++) AL(s.dz_face_new, k, j, J) = m.delta_z[k - 1];
      for (int k = bl; k <= tnew - 1; k++)
        A0(s.dzhalf_face_new, k, j, J) = (AL(s.dz_face_new, k + 1, j, J) + AL(s.dz_face_new, k, j, J)) * 0.5;
      A0(s.dzhalf_face_new, tnew, j, J) = AL(s.dz_face_new, tnew, j, J);
      A0(s.dzhalf_face_new, bl - 1, j, J) = AL(s.dz_face_new, bl, j, J);
    }
  }
  s.top_layer_at_face[j] = tnew;
}

// update_vertical_layers_1.f90:203-254 (nodes)
// update_variables.f90:16-31 for one entity's layer thicknesses: new -> old over the levels
// the layer kernels above write (dz on [bottom, top], dzhalf on [bottom-1, top] of a wet
// entity). Every other entry of the new arrays still holds what the previous update copied,
// so there new and old already agree. One thread per entity walking its levels in chunks of
// four (loads of a chunk issued before its stores); a level row is contiguous across the
// threads of a wave.
__global__ void k_update_dz(double *dz, const double *dz_new, double *dzh, const double *dzh_new,
                            const int32_t *top, const int32_t *bot, int n, int L) {
  int i = gom_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int tp = top[i];
  if (tp == 0) return;
  int b = bot[i];
  for (int k0 = b - 1; k0 <= tp; k0 += 4) {
    double h[4], z[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int k = k0 + q <= tp ? k0 + q : tp;
      h[q] = A0(dzh_new, k, i, n);
      z[q] = AL(dz_new, k >= 1 ? k : 1, i, n);
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int k = k0 + q;
      if (k > tp) break;
      A0(dzh, k, i, n) = h[q];
      if (k >= b) AL(dz, k, i, n) = z[q];
    }
  }
}

__global__ void k_layers_node(DevMesh m, DevState s, DevDims d, double MSL, double dry, int *err) {
  int i = gom_block() * blockDim.x + threadIdx.x;
  const int P = d.P, L = d.L;
  if (i >= P) return;
  int tnew = 0;
  double eta = s.eta_node[i];
  double tw = m.h_node[i] + eta;
  int bl = m.bottom_layer_at_node[i];
  if (tw > dry) {
    double tmp = MSL + eta;
    for (int k = L - 1; k >= 0; k--) {
      if (tmp > m.z_level[k] && tmp <= m.z_level[k + 1]) {
        tnew = k + 1;
        AL(s.dz_node_new, tnew, i, P) = tmp - m.z_level[k];
        break;
      }
    }
    if (tnew < bl) {
      latch(err, GOMGPU_ERR_LAYER_NODE, i + 1);
      return;
    }
    if (tnew == bl) {
      AL(s.dz_node_new, tnew, i, P) = tw;
      A0(s.dzhalf_node_new, tnew, i, P) = tw;
    } else {
      AL(s.dz_node_new, bl, i, P) = m.z_level[bl] - (MSL - m.h_node[i]);
      for (int k = bl + 1; k <= tnew - 1; k++) AL(s.dz_node_new, k, i, P) = m.delta_z[k - 1];
      for (int k = bl; k <= tnew - 1; k++)
        A0(s.dzhalf_node_new, k, i, P) = (AL(s.dz_node_new, k + 1, i, P) + AL(s.dz_node_new, k, i, P)) * 0.5;
    }
  }
  s.top_layer_at_node[i] = tnew;
}

// Element dryness from its nodes and the dz_half checks: update_vertical_layers_1.f90:260-330
__global__ void k_layers_check_cell(DevMesh m, DevState s, DevDims d, double dmin, int *err) {
  int i = gom_block() * blockDim.x + threadIdx.x;
  const int E = d.E;
  if (i >= E) return;
  int N = m.tri_or_quad[i], nsum1 = 0, nsum2 = 0;
  for (int l = 1; l <= N; l++) {
    nsum1 += s.top_layer_at_node[TB(m.nodenum_at_cell, l, i, E) - 1];
    nsum2 += s.top_layer_at_face[TB(m.facenum_at_cell, l, i, E) - 1];
  }
  int t = s.top_layer_at_element[i];
  if (nsum1 == 0) t = 0;
  s.top_layer_at_element[i] = t;
  if (nsum2 == 0 && t != 0) latch(err, GOMGPU_ERR_WET_DRY_SIDES, i + 1);
  for (int k = m.bottom_layer_at_element[i]; k <= t - 1; k++)
    if (A0(s.dzhalf_cell_new, k, i, E) < dmin) latch(err, GOMGPU_ERR_DZHALF_MIN, i + 1);
}

__global__ void k_layers_check_fn(DevMesh m, DevState s, DevDims d, double dmin, int *err) {
  int t = gom_block() * blockDim.x + threadIdx.x;
  const int J = d.J, P = d.P;
  if (t < J) {
    for (int k = m.bottom_layer_at_face[t]; k <= s.top_layer_at_face[t] - 1; k++)
      if (A0(s.dzhalf_face_new, k, t, J) < dmin) latch(err, GOMGPU_ERR_DZHALF_MIN, t + 1);
  } else if (t < J + P) {
    int i = t - J;
    for (int k = m.bottom_layer_at_node[i]; k <= s.top_layer_at_node[i] - 1; k++)
      if (A0(s.dzhalf_node_new, k, i, P) < dmin) latch(err, GOMGPU_ERR_DZHALF_MIN, i + 1);
  }
}

// ================================================================ node velocity
// Face level velocities: calculate_velocity_at_node.f90:31-60
__global__ void k_face_levels(DevMesh m, DevState s, DevDims d) {
  int j = gom_block() * blockDim.x + threadIdx.x;
  const int J = d.J;
  if (j >= J) return;
  int kt = s.top_layer_at_face[j];
  if (kt == 0) return;
  int kb = m.bottom_layer_at_face[j];
  double ct = m.cos_theta[j], st = m.sin_theta[j];
  double un = AL(s.un_face_new, kb, j, J), vn = AL(s.vn_face_new, kb, j, J);
  double uf = un * ct - vn * st, vf = un * st + vn * ct;
  A0(s.u_face_level, kb - 1, j, J) = uf;
  A0(s.v_face_level, kb - 1, j, J) = vf;
  for (int k = kb; k <= kt - 1; k++) {
    double un1 = AL(s.un_face_new, k + 1, j, J), vn1 = AL(s.vn_face_new, k + 1, j, J);
    double uf1 = un1 * ct - vn1 * st, vf1 = un1 * st + vn1 * ct;
    double r = AL(s.dz_face, k, j, J) * 0.5 / A0(s.dzhalf_face, k, j, J);
    A0(s.u_face_level, k, j, J) = uf + r * (uf1 - uf);
    A0(s.v_face_level, k, j, J) = vf + r * (vf1 - vf);
    uf = uf1;
    vf = vf1;
  }
  A0(s.u_face_level, kt, j, J) = uf;
  A0(s.v_face_level, kt, j, J) = vf;
}

// Node u, v (calculate_velocity_at_node.f90:67-115) and w (:120-138); one thread per
// (level, node). The node's faces and cells come from the static stencil tables (one
// coalesced load each instead of the reference's cell -> face lookups), visited in the
// reference's order; the weight sum is precomputed in that order.
#define NV_CHUNK 2
#define NV_LEVELS GOM_NV_LEVELS
__global__ void GOM_WAVES(GOM_WAVES_NODE_VEL) k_node_vel(DevMesh m, DevState s, DevDims d) {
  // NV_LEVELS consecutive levels of one node per thread: the node's stencil entries are
  // loaded once per chunk of adjacent cells and serve every level of the thread
  const int P = d.P, E = d.E, J = d.J, L1 = d.L + 1;
  const int nlc = (L1 + NV_LEVELS - 1) / NV_LEVELS;
  int t = gom_block() * blockDim.x + threadIdx.x;
  if (t >= P * nlc) return;
  const int i = t % P, k0 = (t / P) * NV_LEVELS;
  const int nlev = L1 - k0 < NV_LEVELS ? L1 - k0 : NV_LEVELS;
  const int nadj = m.adj_cells_at_node[i];
  double u[NV_LEVELS], v[NV_LEVELS], w[NV_LEVELS], tw[NV_LEVELS];
  int icount[NV_LEVELS];
#pragma unroll
  for (int z = 0; z < NV_LEVELS; z++) {
    u[z] = v[z] = w[z] = tw[z] = 0.0;
    icount[z] = 0;
  }
  // the adjacent cells in chunks of NV_CHUNK: each chunk's stencil entries and the wet tops
  // they index are loaded once, then per level the face and cell values, and the sums run
  // in the reference's order (slots beyond nadj are not added)
  for (int j0 = 0; j0 < nadj; j0 += NV_CHUNK) {
    int ifn[2 * NV_CHUNK], isn[NV_CHUNK], ftop[2 * NV_CHUNK], ctop[NV_CHUNK], fkb[2 * NV_CHUNK], ckb[NV_CHUNK];
    double lf[2 * NV_CHUNK], fw[2 * NV_CHUNK], car[NV_CHUNK];
#pragma unroll
    for (int q = 0; q < NV_CHUNK; q++) {
      const int jj = j0 + q < nadj ? j0 + q : j0;
      const size_t oc = (size_t)jj * P + i;
      const int4 ff = *(const int4 *)(m.ns_face + 4 * oc);
      const double2 ll = *(const double2 *)(m.ns_Lf + 4 * oc), ww = *(const double2 *)(m.ns_Lf + 4 * oc + 2);
      const int2 cc = *(const int2 *)(m.ns_cell + 2 * oc);
      ifn[2 * q] = ff.x;
      ifn[2 * q + 1] = ff.y;
      fkb[2 * q] = ff.z;
      fkb[2 * q + 1] = ff.w;
      lf[2 * q] = ll.x;
      lf[2 * q + 1] = ll.y;
      fw[2 * q] = ww.x;
      fw[2 * q + 1] = ww.y;
      isn[q] = cc.x;
      ckb[q] = cc.y;
      car[q] = m.ns_area[oc];
    }
#pragma unroll
    for (int q = 0; q < NV_CHUNK; q++) {
      ftop[2 * q] = s.top_layer_at_face[ifn[2 * q]];
      ftop[2 * q + 1] = s.top_layer_at_face[ifn[2 * q + 1]];
      ctop[q] = s.top_layer_at_element[isn[q]];
    }
#pragma unroll
    for (int z = 0; z < NV_LEVELS; z++) {
      const int k = k0 + (z < nlev ? z : 0);   // levels past the top of the range repeat k0
      double fu[2 * NV_CHUNK], fv[2 * NV_CHUNK], cw[NV_CHUNK];
#pragma unroll
      for (int f = 0; f < 2 * NV_CHUNK; f++) {
        const int kvt = k < ftop[f] ? k : ftop[f];
        fu[f] = A0(s.u_face_level, kvt, ifn[f], J);
        fv[f] = A0(s.v_face_level, kvt, ifn[f], J);
      }
#pragma unroll
      for (int q = 0; q < NV_CHUNK; q++) {
        const int kvt = k < ctop[q] ? k : ctop[q];
        cw[q] = A0(s.wn_cell, kvt, isn[q], E);
      }
#pragma unroll
      for (int q = 0; q < NV_CHUNK; q++) {
        if (j0 + q >= nadj) continue;
#pragma unroll
        for (int h = 0; h < 2; h++) {
          const int f = 2 * q + h;
          if (ftop[f] != 0 && k >= fkb[f] - 1) {
            u[z] = u[z] + fu[f] / lf[f] * fw[f];
            v[z] = v[z] + fv[f] / lf[f] * fw[f];
            icount[z]++;
          }
        }
        if (ctop[q] != 0 && k >= ckb[q] - 1) {
          tw[z] = tw[z] + car[q];
          w[z] = w[z] + cw[q] * car[q];
        }
      }
    }
  }
  const double weight = m.ns_weight[i];
#pragma unroll
  for (int z = 0; z < NV_LEVELS; z++) {
    if (z >= nlev) continue;
    const int k = k0 + z;
    double uu = u[z], vv = v[z], ww = w[z];
    if (icount[z] != 0) {
      uu = uu / weight;
      vv = vv / weight;
    }
    if (tw[z] != 0.0) ww = ww / tw[z];
    A0(s.u_node, k, i, P) = uu;
    A0(s.v_node, k, i, P) = vv;
    A0(s.w_node, k, i, P) = ww;
  }
}

// River face velocity from the discharge (compute_boundary_velocity.f90:32-46):
// Qu = -Q_add / (face_length * (h_face + mean eta of the face nodes)).
__global__ void k_river_velocity(DevMesh m, DevState s, DevForcing f, DevDims d, double *qu) {
  int i = gom_block() * blockDim.x + threadIdx.x;
  const int J = d.J;
  if (i >= d.NQB) return;
  // the river face is the face of the river cell flagged with this river id
  int e = m.Qb_cell_id[i] - 1;
  int N = m.tri_or_quad[e];
  for (int l = 1; l <= N; l++) {
    int j = TB(m.facenum_at_cell, l, e, d.E) - 1;
    if (m.isflowside3[j] == i + 1) {
      int n1 = TB(m.nodenum_at_face, 1, j, J) - 1, n2 = TB(m.nodenum_at_face, 2, j, J) - 1;
      double depth = m.h_face[j] + (s.eta_node[n1] + s.eta_node[n2]) * 0.5;
      qu[i] = -f.Q_add[i] / (m.face_length[j] * depth);
      return;
    }
  }
}

// ---- hurricane fields and wind stress (read_hurricane_ser_1.f90, calculate_wind_stress.f90)
// Node fields of one step. Phase 2 (every hurricane_read_interval steps) moves the later
// snapshot into the earlier pair and the interpolated field, then takes the next snapshot
// (read_hurricane_ser_1.f90:18-46); phase 1 adds one interval's increment to the
// interpolated field (read_hurricane_ser_1.f90:48-60), so the field is the running sum the
// reference accumulates; phase 0 (outside the hurricane window) leaves everything.
// The next hurricane snapshot of a Holland cyclone (gomgpu_holland), at a step whose phase
// takes it (scal[5] == 2), for the centre (scal[3], scal[7]); one thread per node. The
// arithmetic is the host generator's (meshgen.HollandTrack.at), operation for operation.
__global__ void k_holland_next(DevMesh m, DevForcing f, gomgpu_holland h, int P) {
  const int i = gom_block() * blockDim.x + threadIdx.x;
  if (i >= P || f.scal[5] != 2.0) return;
  const double cx = f.scal[3], cy = f.scal[7];
  const double dp = (h.p_env - h.p_min) * 100.0;
  const double dx = m.x_node[i] - cx, dy = m.y_node[i] - cy;
  const double r = hypot(dx, dy);
  const double rs = r > 1.0 ? r : 1.0;
  const double s = pow(h.r_max / rs, h.b);
  const double es = exp(-s);
  const double p = h.p_min + (h.p_env - h.p_min) * es;
  const double rf2 = rs * h.coriolis / 2.0;
  const double vg = sqrt(h.b * dp * s * es / h.rho_air + rf2 * rf2) - rs * h.coriolis / 2.0;
  const double vs = r > 1.0 ? h.surface_factor * vg : 0.0;
  const_cast<double *>(f.wind_u_next)[i] = -vs * dy / rs + h.vx;
  const_cast<double *>(f.wind_v_next)[i] = vs * dx / rs + h.vy;
  const_cast<double *>(f.air_p_next)[i] = p * 100.0;
}

__global__ void k_hurricane_nodes(DevForcing f, int P, double ninterval) {
  int i = gom_block() * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const double phase = f.scal[5];
  double *u0 = const_cast<double *>(f.wind_u0), *v0 = const_cast<double *>(f.wind_v0);
  double *p0 = const_cast<double *>(f.air_p0);
  double *u1 = const_cast<double *>(f.wind_u1), *v1 = const_cast<double *>(f.wind_v1);
  double *p1 = const_cast<double *>(f.air_p1);
  double *u2 = const_cast<double *>(f.wind_u2), *v2 = const_cast<double *>(f.wind_v2);
  double *p2 = const_cast<double *>(f.air_p2);
  if (phase == 2.0) {
    const double a = u2[i], b = v2[i], c = p2[i];
    u0[i] = a;
    v0[i] = b;
    p0[i] = c;
    u1[i] = a;
    v1[i] = b;
    p1[i] = c;
    u2[i] = f.wind_u_next[i];
    v2[i] = f.wind_v_next[i];
    p2[i] = f.air_p_next[i];
  } else if (phase == 1.0) {
    u1[i] = u1[i] + (u2[i] - u0[i]) / ninterval;
    v1[i] = v1[i] + (v2[i] - v0[i]) / ninterval;
    p1[i] = p1[i] + (p2[i] - p0[i]) / ninterval;
  }
}

// Face winds (the mean of the face's two node winds, inside the hurricane window) and the
// wind stress of every face (calculate_wind_stress.f90:88-114): drag coefficient of the
// chosen formula capped at 0.003 (a default-real literal in the reference), stress along
// the face normal and tangent, times the wind spin-up factor.
__global__ void k_wind_faces(DevMesh m, DevForcing f, int J, double rho_ratio, int formula) {
  int j = gom_block() * blockDim.x + threadIdx.x;
  if (j >= J) return;
  const double phase = f.scal[5], spin = f.scal[4];
  double *uf_ = const_cast<double *>(f.wind_u_at_face), *vf_ = const_cast<double *>(f.wind_v_at_face);
  double uf, vf;
  if (phase != 0.0) {
    int n1 = TB(m.nodenum_at_face, 1, j, J) - 1, n2 = TB(m.nodenum_at_face, 2, j, J) - 1;
    uf = (f.wind_u1[n1] + f.wind_u1[n2]) * 0.5;
    vf = (f.wind_v1[n1] + f.wind_v1[n2]) * 0.5;
    uf_[j] = uf;
    vf_[j] = vf;
  } else {
    uf = uf_[j];
    vf = vf_[j];
  }
  const double c = m.cos_theta[j], s = m.sin_theta[j];
  const double ws = sqrt(uf * uf + vf * vf);
  const double wn = uf * c + vf * s;
  const double wt = -(uf * s) + vf * c;
  double a = 0.61, b = 0.063;
  if (formula == 1) {
    a = 0.75;
    b = 0.067;
  } else if (formula == 3) {
    a = 0.80;
    b = 0.065;
  }
  double cda = 0.001 * (a + b * ws);
  const double cap = (double)0.003f;
  cda = cap < cda ? cap : cda;
  const double g = cda * rho_ratio * ws;
  const_cast<double *>(f.Gamma_T)[j] = g;
  const_cast<double *>(f.wind_stress_normal)[j] = g * wn * spin;
  const_cast<double *>(f.wind_stress_tangnt)[j] = g * wt * spin;
}
