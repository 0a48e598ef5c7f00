"""Host-side set-up of a model from a generated case: connectivity and geometry, boundary
lists, vertical layers and the cold-start state -- what the reference computes once before
its time loop (GOM_Rv1.1.0/source: read_node_inp.f90, read_cell_inp.f90,
set_geometry_1.f90, set_geometry_2.f90, read_main_inp.f90, calculate_coriolis_factor.f90,
update_vertical_layers_0.f90, setup_cold_start.f90, read_salt_init.f90).

Arithmetic follows the reference expression by expression (numpy for the elementwise
+,-,*,/,sqrt, which are correctly rounded, and Python's math for atan2/cos/sin, which call
the same libm as the reference), so the arrays it produces are bit-identical to the
reference's; the tests check this against the reference's own geometry dump.
The equation of state of the initial state is evaluated by the device EOS stage.
"""
from __future__ import annotations

import math
from typing import Tuple

import numpy as np

from . import abi
from .gomio import Case

PI = 3.141592653589793
DEG2RAD = PI / 180.0


def _sen(N, l, w):
    v = l + w
    return np.where(v > N, v - N, v)


def params_for(case: Case, precision: int = 0) -> "abi.Params":
    z = np.asarray(case.z_level, dtype=np.float64)
    L = case.maxlayer
    dz = z[1:] - z[:-1]
    ntr = (1 if case.is_salt else 0) + (1 if case.is_temp else 0)
    tid = ([1] if case.is_salt else []) + ([2] if case.is_temp else [])
    if case.transport_flag == 1 and case.transport_solver == 2:
        # the ELM solver carries salt and temperature both (solve_transport_equation_ELM_v8.f90)
        ntr, tid = 2, [1, 2]
    tid += [0, 0]
    d = dict(
        maxnod=case.maxnod, maxele=case.maxele, maxface=0, maxlayer=L,
        max_no_neighbor_node=int(case.max_no_neighbor_node),
        num_ob_cell=len(case.open_boundaries), num_Qb_cell=len(case.rivers),
        maxtran2=ntr if case.transport_flag == 1 else 0, tran_id1=tid[0], tran_id2=tid[1],
        advection_flag=case.advection_flag, ELM_backtrace_flag=case.elm_backtrace_flag,
        ELM_min_iter=case.elm_min_iter, ELM_max_iter=case.elm_max_iter, bf_flag=case.bf_flag,
        baroclinic_flag=case.baroclinic_flag, ana_density=case.ana_density, i_density_flag=0,
        transport_flag=case.transport_flag,
        # the ELM solver reads no TVD card (C7_1), whose values stay zero
        trans_sub_iter=case.trans_sub_iter if case.transport_solver != 2 else 0,
        h_flux_limiter=case.h_flux_limiter if case.transport_solver != 2 else 0,
        v_flux_limiter=case.v_flux_limiter if case.transport_solver != 2 else 0,
        heat_option=1 if case.is_temp else 0, max_iteration_pcg=case.max_iteration_pcg,
        precision=precision,
        dt=case.dt, theta=case.theta, gravity=float(case.gravity),
        rho_o=float(case.rho_o), MSL=case.msl,
        dry_depth=case.dry_depth, adv_onoff_depth=case.adv_onoff_depth,
        smagorinsky_parameter=case.smagorinsky, Ah_0=case.ah0, Kh_0=case.kh0,
        von_Karman=case.von_karman, error_tolerance=case.error_tolerance,
        # read_main_inp.f90:262, with (0.5 - 1.0e-4) folded in single precision
        denominator_min_for_matrix=min(case.dry_depth * float(np.float32(0.5) - np.float32(1.0e-4)),
                                       float(dz.min()) * float(np.float32(0.5) - np.float32(1.0e-4))),
        delta_z_min=float(dz.min()),
        # C30/C31 (read_hurricane_ser_0.f90:29: hurricane_read_interval = int(hurricane_dt/dt))
        hurricane_flag=int(case.hurricane_flag), wind_formular=int(case.wind_formular),
        hurricane_read_interval=int(case.hurricane_dt / case.dt) if case.hurricane_flag == 1 else 0,
        transport_solver=int(case.transport_solver) if case.transport_flag == 1 else 1,
        rho_a=float(case.rho_a))
    d.update(turbulence_params(case))
    return abi.params_from_dict(d)


# Second-equation constants of the generic length scale closure (Umlauf & Burchard 2003,
# J. Mar. Res. 61, Table 1; MY and k-kl as in Mellor & Yamada 1982): psi = cmiu0^p k^m l^n.
GLS_MODELS = {
    "my": dict(rpub=0.0, rmub=1.0, rnub=1.0, cpsi1=0.9, cpsi2=0.5, schk=2.44, schpsi=2.44),
    "kl": dict(rpub=0.0, rmub=1.0, rnub=1.0, cpsi1=0.9, cpsi2=0.5, schk=2.44, schpsi=2.44),
    "ke": dict(rpub=3.0, rmub=1.5, rnub=-1.0, cpsi1=1.44, cpsi2=1.92, schk=1.0, schpsi=1.3),
    "kw": dict(rpub=-1.0, rmub=0.5, rnub=-1.0, cpsi1=0.555, cpsi2=0.833, schk=2.0, schpsi=2.0),
    "ub": dict(rpub=2.0, rmub=1.0, rnub=-0.67, cpsi1=1.0, cpsi2=1.22, schk=0.8, schpsi=1.07),
}
GLS_IDS = {"my": 0, "kl": 1, "ke": 2, "kw": 3, "ub": 4}
TURBULENCE_PARAMS = ("turbulence_flag", "turbulence_model", "stability_function", "q2min", "psimin",
                     "eps_min", "bgdiff", "diffmax", "xlmin1", "xlmin2", "cmiu0", "cpsi1", "cpsi2",
                     "rpub", "rmub", "rnub", "schk", "schpsi", "h1_pp", "h2_pp", "vdmax_pp1",
                     "vdmax_pp2", "vdmin_pp1", "vdmin_pp2", "tdmin_pp")


def turbulence_params(case: Case) -> dict:
    """Closure parameters of the ABI (the module variables of turbulence_closure_v1.f90).

    cmiu0 = (2^1.5 / B1)^(1/3) with B1 = 16.6 makes the dissipation cmiu0^3 k^1.5 / l equal
    to Mellor-Yamada's q^3 / (B1 l) with q^2 = 2 k."""
    mdl = GLS_MODELS[case.turbulence_model]
    cmiu0 = (2.0 ** 1.5 / 16.6) ** (1.0 / 3.0)
    psimin = cmiu0 ** mdl["rpub"] * case.q2min ** mdl["rmub"] * case.xlmin ** mdl["rnub"]
    return dict(
        turbulence_flag=case.turbulence_flag, turbulence_model=GLS_IDS[case.turbulence_model],
        stability_function={"GA": 0, "KC": 1}[case.stability_function],
        q2min=case.q2min, psimin=psimin, eps_min=case.eps_min, bgdiff=case.bgdiff,
        diffmax=case.diffmax, xlmin1=case.xlmin, xlmin2=case.xlmin, cmiu0=cmiu0,
        h1_pp=case.pp_depths[0], h2_pp=case.pp_depths[1], vdmax_pp1=case.pp_vdmax[0],
        vdmax_pp2=case.pp_vdmax[1], vdmin_pp1=case.pp_vdmin[0], vdmin_pp2=case.pp_vdmin[1],
        tdmin_pp=case.pp_tdmin, **mdl)


def build(case: Case, precision: int = 0) -> Tuple["abi.Params", "abi.Arrays", "abi.Arrays"]:
    """Return (params, mesh, state) in the ABI layout for a generated case."""
    x = np.asarray(case.x, dtype=np.float64)
    y = np.asarray(case.y, dtype=np.float64)
    h_node = np.asarray(case.depth, dtype=np.float64) + float(case.h_node_adjust)   # read_node_inp.f90:64
    P = x.shape[0]
    cells = np.asarray(case.cells, dtype=np.int64)
    E = cells.shape[0]
    N = np.where(cells[:, 3] != 0, 4, 3)
    L = case.maxlayer
    NB = int(case.max_no_neighbor_node)
    nc = cells.T.copy()                         # (4, E), 1-based node ids

    # ---- area (read_cell_inp.f90:67-80)
    def carea(x1, x2, x3, y1, y2, y3):
        return np.abs((x1 - x3) * (y2 - y3) - (x2 - x3) * (y1 - y3)) / 2.0
    X = lambda l: x[nc[l] - 1]  # noqa: E731
    Y = lambda l: y[nc[l] - 1]  # noqa: E731
    a123 = carea(X(0), X(1), X(2), Y(0), Y(1), Y(2))
    n4 = np.maximum(nc[3], 1)
    a134 = carea(X(0), X(2), x[n4 - 1], Y(0), Y(2), y[n4 - 1])
    area = np.where(N == 4, a123 + a134, a123)

    # ---- cells around nodes (set_geometry_1.f90:109-127)
    cid = np.repeat(np.arange(1, E + 1), 4)
    slot = np.tile(np.arange(1, 5), E)
    nod = nc.T.ravel()
    ok = nod != 0
    cid, slot, nod = cid[ok], slot[ok], nod[ok]
    order = np.lexsort((slot, cid, nod))        # by node, then cell, then slot
    cid_s, slot_s, nod_s = cid[order], slot[order], nod[order]
    adj_cells_at_node = np.bincount(nod_s, minlength=P + 1)[1:].astype(np.int32)
    if adj_cells_at_node.max() > NB or adj_cells_at_node.min() == 0:
        raise ValueError("node with too many or no adjacent cells")
    start = np.concatenate([[0], np.cumsum(adj_cells_at_node)[:-1]])
    rank = np.arange(nod_s.size) - start[nod_s - 1]
    adj_cellnum_at_node = np.zeros((NB, P), np.int32)
    node_count = np.zeros((NB, P), np.int32)
    adj_cellnum_at_node[rank, nod_s - 1] = cid_s
    node_count[rank, nod_s - 1] = slot_s

    # ---- nodes around nodes (set_geometry_1.f90:160-186)
    adj_nodes_at_node = np.zeros(P, np.int32)
    adj_nodenum_at_node = np.zeros((NB + 1, P), np.int32)
    cand_node, cand_nb, cand_pos = [], [], []
    for r in range(NB):
        has = r < adj_cells_at_node
        nodes_i = np.flatnonzero(has) + 1
        cells_r = adj_cellnum_at_node[r, has]
        for l in range(4):
            nb = nc[l, cells_r - 1]
            cand_node.append(nodes_i)
            cand_nb.append(nb)
            cand_pos.append(np.full(nodes_i.size, r * 4 + l))
    cn = np.concatenate(cand_node)
    cb = np.concatenate(cand_nb)
    cp = np.concatenate(cand_pos)
    # The reference tests `nd == i` inside a loop over the neighbours found so far, which
    # is empty for the first candidate: a node whose first adjacent cell starts with the
    # node itself lists itself as its first neighbour (set_geometry_1.f90:169-183).
    keep = (cb != 0) & ((cb != cn) | (cp == 0))
    cn, cb, cp = cn[keep], cb[keep], cp[keep]
    o = np.lexsort((cp, cb, cn))
    cn, cb, cp = cn[o], cb[o], cp[o]
    first = np.ones(cn.size, bool)
    first[1:] = (cn[1:] != cn[:-1]) | (cb[1:] != cb[:-1])
    cn, cb, cp = cn[first], cb[first], cp[first]
    o = np.lexsort((cp, cn))
    cn, cb = cn[o], cb[o]
    cnt = np.bincount(cn, minlength=P + 1)[1:]
    if cnt.max() > NB:
        raise ValueError("node with too many neighbour nodes")
    adj_nodes_at_node[:] = cnt
    st2 = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    rk = np.arange(cn.size) - st2[cn - 1]
    adj_nodenum_at_node[rk + 1, cn - 1] = cb

    # ---- neighbour cells across faces (set_geometry_1.f90:206-226)
    l_idx = np.arange(1, 5)[:, None]
    nd1 = np.take_along_axis(nc, _sen(N[None, :], l_idx, 1) - 1, axis=0)
    nd2 = np.take_along_axis(nc, _sen(N[None, :], l_idx, 2) - 1, axis=0)
    valid = l_idx <= N[None, :]
    ekey = np.where(valid, np.minimum(nd1, nd2) * (P + 1) + np.maximum(nd1, nd2), -1)
    flat_key = ekey.T.ravel()
    flat_cell = np.repeat(np.arange(1, E + 1), 4)
    flat_slot = np.tile(np.arange(1, 5), E)
    vm = flat_key >= 0
    fk, fc, fs = flat_key[vm], flat_cell[vm], flat_slot[vm]
    o = np.argsort(fk, kind="stable")
    fk, fc, fs = fk[o], fc[o], fs[o]
    same_next = np.zeros(fk.size, bool)
    same_next[:-1] = fk[1:] == fk[:-1]
    if np.any(same_next[:-1] & same_next[1:]):
        raise ValueError("an edge is shared by more than two cells")
    adj_cell = np.zeros((4, E), np.int32)
    a_idx = np.flatnonzero(same_next)
    adj_cell[fs[a_idx] - 1, fc[a_idx] - 1] = fc[a_idx + 1]
    adj_cell[fs[a_idx + 1] - 1, fc[a_idx + 1] - 1] = fc[a_idx]

    # ---- faces (set_geometry_1.f90:241-346)
    cells_all = np.repeat(np.arange(1, E + 1), 4)
    slots_all = np.tile(np.arange(1, 5), E)
    adj_all = adj_cell.T.ravel()
    validf = (slots_all <= np.repeat(N, 4))
    newf = validf & ((adj_all == 0) | (cells_all < adj_all))
    J = int(newf.sum())
    fcell = cells_all[newf]
    fslot = slots_all[newf]
    fid = np.arange(1, J + 1)
    facenum = np.zeros((4, E), np.int32)
    facenum[fslot - 1, fcell - 1] = fid
    fnb = adj_all[newf]
    n1f = nd1[fslot - 1, fcell - 1]
    n2f = nd2[fslot - 1, fcell - 1]
    # the neighbour's slot that points back (first match)
    has_nb = fnb > 0
    back = np.zeros(J, np.int64)
    for l in range(4, 0, -1):
        hit = has_nb & (adj_cell[l - 1, np.maximum(fnb, 1) - 1] == fcell)
        back = np.where(hit, l, back)
    facenum[back[has_nb] - 1, fnb[has_nb] - 1] = fid[has_nb]
    xn1, yn1, xn2, yn2 = x[n1f - 1], y[n1f - 1], x[n2f - 1], y[n2f - 1]
    x_face = (xn1 + xn2) * 0.5
    y_face = (yn1 + yn2) * 0.5
    h_face = (h_node[n1f - 1] + h_node[n2f - 1]) * 0.5
    face_length = np.sqrt((xn2 - xn1) ** 2 + (yn2 - yn1) ** 2)
    adj_face = np.stack([fcell, fnb]).astype(np.int32)
    node_face = np.stack([n1f, n2f]).astype(np.int32)

    # ---- sign, centres, depth (set_geometry_1.f90:375-419)
    fnum = facenum.astype(np.int64)
    a1 = adj_face[0, np.maximum(fnum, 1) - 1].astype(np.int64)
    a2 = adj_face[1, np.maximum(fnum, 1) - 1].astype(np.int64)
    ii = np.arange(1, E + 1)[None, :]
    num = a2 - 2 * ii + a1
    den = a2 - a1
    sign = np.where(fnum > 0, np.trunc(num / np.where(den == 0, 1, den)), 0).astype(np.int32)
    x_cell = np.zeros(E)
    y_cell = np.zeros(E)
    h_cell = np.zeros(E)
    for l in range(4):
        m_ = (l + 1) <= N
        x_cell = np.where(m_, x_cell + x[np.maximum(nc[l], 1) - 1], x_cell)
        y_cell = np.where(m_, y_cell + y[np.maximum(nc[l], 1) - 1], y_cell)
        hf = h_face[np.maximum(fnum[l], 1) - 1]
        h_cell = np.where(m_ & (hf > h_cell), hf, h_cell)
    x_cell = x_cell / N
    y_cell = y_cell / N

    # ---- delta_j and angles (set_geometry_1.f90:446-531), Voronoi off
    inter = fnb > 0
    c1 = fcell - 1
    c2 = np.maximum(fnb, 1) - 1
    dj_int = np.sqrt((x_cell[c2] - x_cell[c1]) ** 2 + (y_cell[c2] - y_cell[c1]) ** 2)
    dj_bnd = np.sqrt((x_cell[c1] - x_face) ** 2 + (y_cell[c1] - y_face) ** 2) * 2.0
    delta_j = np.where(inter, dj_int, dj_bnd)
    # the angles through the C library's atan2/cos/sin (math), as the reference's libm does;
    # numpy's vectorised versions may differ in the last bit
    def libm(f, *args):
        return np.fromiter(map(f, *(a.tolist() for a in args)), np.float64, J)
    th0 = libm(math.atan2, xn1 - xn2, yn2 - yn1)
    cos_t = libm(math.cos, th0)
    sin_t = libm(math.sin, th0)
    th1 = np.where(inter, libm(math.atan2, y_cell[c2] - y_cell[c1], x_cell[c2] - x_cell[c1]), th0)
    th2 = th1 - th0
    cos_t2 = libm(math.cos, th2)
    sin_t2 = libm(math.sin, th2)

    # ---- boundaries (set_geometry_2.f90:20-82, read_main_inp.f90:841-887)
    pkey = n1f.astype(np.int64) * (P + 1) + n2f
    porder = np.argsort(pkey, kind="stable")   # equal keys keep face order: the first wins
    pkey_sorted = pkey[porder]

    def find_cell_face(a, b):
        key = int(a) * (P + 1) + int(b)
        at = int(np.searchsorted(pkey_sorted, key))
        if at >= J or pkey_sorted[at] != key:
            raise ValueError(f"boundary nodes {a},{b} are not a counter-clockwise cell edge")
        j = int(porder[at]) + 1
        return int(fcell[j - 1]), j
    nob = len(case.open_boundaries)
    ob_cell_id = np.zeros(max(nob, 1), np.int32)
    ob_element_flag = np.zeros(E, np.int32)
    btype = np.zeros(J, np.int32)
    for i, b in enumerate(case.open_boundaries, 1):
        c, j = find_cell_face(b.n1, b.n2)
        ob_cell_id[i - 1] = c
        ob_element_flag[c - 1] = i
        btype[j - 1] = i
    btype = np.where((fnb == 0) & (btype == 0), -1, btype).astype(np.int32)
    nqb = len(case.rivers)
    qb_cell = np.zeros(max(nqb, 1), np.int32)
    qb_flag = np.zeros(E, np.int32)
    isflow3 = np.zeros(J, np.int32)
    for i, r in enumerate(case.rivers, 1):
        c, j = find_cell_face(r.n1, r.n2)
        qb_cell[i - 1] = c
        isflow3[j - 1] = i
        qb_flag[c - 1] = i

    # ---- bottom layers (set_geometry_2.f90:101-189)
    z = np.asarray(case.z_level, dtype=np.float64)
    MSL = case.msl

    def bottom(h):
        lev = MSL - h
        bl = np.zeros(h.shape[0], np.int32)
        for k in range(L - 1, -1, -1):
            bl = np.where((lev >= z[k]) & (lev < z[k + 1]), k + 1, bl)
        if np.any(bl == 0):
            raise ValueError("depth outside the vertical grid")
        return bl
    bl_node, bl_cell, bl_face = bottom(h_node), bottom(h_cell), bottom(h_face)

    p = params_for(case, precision)
    p.maxface = J
    mesh = abi.mesh_arrays(p)
    omega = float(np.float32(7.292e-5))
    cor = 2 * omega * math.sin(case.lat * DEG2RAD) if case.coriolis_option == 1 else 0.0
    if case.coriolis_option not in (0, 1):
        raise NotImplementedError("only Coriolis options 0 and 1")
    vals = dict(x_node=x, y_node=y, h_node=h_node, x_cell=x_cell, y_cell=y_cell, h_cell=h_cell,
                area=area, face_length=face_length, x_face=x_face, y_face=y_face, h_face=h_face,
                delta_j=delta_j, cos_theta=cos_t, sin_theta=sin_t, cos_theta2=cos_t2, sin_theta2=sin_t2,
                coriolis_factor=np.full(J, cor),
                # read_main_inp.f90:226-244: only the coefficient of the chosen law is set
                manning=np.full(J, case.manning if case.bf_flag == 1 else 0.0),
                bf_height=np.full(J, case.bf_height if case.bf_flag == 2 else 0.0),
                z_level=z, delta_z=z[1:] - z[:-1],
                tri_or_quad=N.astype(np.int32), nodenum_at_cell=nc.astype(np.int32),
                adj_cellnum_at_cell=adj_cell, facenum_at_cell=facenum, sign_in_outflow=sign,
                adj_cellnum_at_face=adj_face, nodenum_at_face=node_face,
                adj_cells_at_node=adj_cells_at_node, adj_cellnum_at_node=adj_cellnum_at_node,
                node_count_each_element=node_count, adj_nodes_at_node=adj_nodes_at_node,
                adj_nodenum_at_node=adj_nodenum_at_node, boundary_type_of_face=btype,
                isflowside3=isflow3, ob_element_flag=ob_element_flag, Qb_element_flag=qb_flag,
                bottom_layer_at_node=bl_node, bottom_layer_at_element=bl_cell,
                bottom_layer_at_face=bl_face,
                initial_wetdry_node=(h_node <= 0.0).astype(np.int32),
                ob_cell_id=ob_cell_id, Qb_cell_id=qb_cell)
    for k, v in vals.items():
        mesh[k] = v
    state = hot_start(case, p, mesh) if case.restart_in == 1 else cold_start(case, p, mesh)
    return p, mesh, state


def _layers0(h, eta, bl, z, L, MSL, dry, kind):
    """update_vertical_layers_0.f90 for one entity class; returns (top, dz, dzhalf).

    Whole-array form of the reference's per-entity statements (each value is the same
    expression, so the result is the same bit for bit), in the ABI's layer-fastest order."""
    n = h.shape[0]
    tw = h + eta
    tmp = MSL + eta
    wet = tw > dry
    # top layer: the k with z[k] < MSL + eta <= z[k+1] (z increases)
    kk = np.searchsorted(z, tmp, side="left") - 1
    top = np.where(wet & (kk >= 0) & (kk <= L - 1), kk + 1, 0).astype(np.int32)
    if np.any(wet & (top < bl)):
        raise ValueError("initial surface below the bottom layer")
    one = wet & (top == bl)
    many = wet & (top > bl)
    idx = np.arange(n)
    # interior layers of the deeper columns, then the top and bottom layers (disjoint rows)
    K = np.arange(1, L + 1)[:, None]
    dz = np.where(many & (K > bl) & (K < top), (z[1:] - z[:-1])[:, None], 0.0)
    w, o, mm = idx[wet], idx[one], idx[many]
    dz[top[w] - 1, w] = tmp[w] - z[top[w] - 1]
    dz[top[o] - 1, o] = tw[o]
    dz[bl[mm] - 1, mm] = z[bl[mm]] - (MSL - h[mm])
    dzh = np.zeros((L + 1, n))
    if L > 1:
        K1 = np.arange(1, L)[:, None]
        dzh[1:L] = np.where(many & (K1 >= bl) & (K1 <= top - 1), (dz[1:L] + dz[0:L - 1]) * 0.5, 0.0)
    if kind == "face":
        dzh[top[o] - 1, o] = tw[o]
    dzh[top[o], o] = tw[o]
    if kind in ("face", "node"):
        dzh[top[mm], mm] = dz[top[mm] - 1, mm]
        dzh[bl[mm] - 1, mm] = dz[bl[mm] - 1, mm]
    return top, np.asfortranarray(dz), np.asfortranarray(dzh)


def cold_start(case: Case, p: "abi.Params", mesh: "abi.Arrays") -> "abi.Arrays":
    """Initial state after prepare_gom with a cold start (EOS excluded)."""
    s = abi.state_arrays(p)
    L, P, E, J = p.maxlayer, p.maxnod, p.maxele, p.maxface
    z = mesh["z_level"]
    MSL, dry = p.MSL, p.dry_depth
    tc, dzc, dzhc = _layers0(mesh["h_cell"], np.zeros(E), mesh["bottom_layer_at_element"], z, L, MSL, dry, "cell")
    tf, dzf, dzhf = _layers0(mesh["h_face"], np.zeros(J), mesh["bottom_layer_at_face"], z, L, MSL, dry, "face")
    tn, dzn, dzhn = _layers0(mesh["h_node"], np.zeros(P), mesh["bottom_layer_at_node"], z, L, MSL, dry, "node")
    nc = mesh.view2("nodenum_at_cell")
    fc = mesh.view2("facenum_at_cell")
    N = mesh["tri_or_quad"]
    nsum1 = np.zeros(E, np.int64)
    for l in range(4):
        m_ = l < N
        nsum1 += np.where(m_, tn[np.maximum(nc[l], 1) - 1], 0)
    tc = np.where(nsum1 == 0, 0, tc).astype(np.int32)
    s["top_layer_at_element"] = tc
    s["top_layer_at_face"] = tf
    s["top_layer_at_node"] = tn
    s["dz_cell"] = dzc
    s["dzhalf_cell"] = dzhc
    s["dz_face"] = dzf
    s["dzhalf_face"] = dzhf
    s["dz_node"] = dzn
    s["dzhalf_node"] = dzhn
    nsub = {0: case.elm_sub_iter, 1: 1}.get(case.elm_backtrace_flag, 0)
    s["num_sub_elm_iteration"] = np.full(L * P, nsub, np.int32)
    av = np.zeros((L + 1, J), order="F")
    av[1:, :] = case.av0
    s["Av"] = av
    kv = np.zeros((L + 1, E), order="F")
    kv[1:, :] = case.kv0
    s["Kv"] = kv
    # two-equation closure state: k at its floor, length scale at its minimum
    s["q2"] = np.full((L + 1, J), case.q2min, order="F")
    s["xl"] = np.full((L + 1, J), case.xlmin, order="F")
    # tracers (allocate_variables.f90:...: salt 0, temp 20; read_salt_init.f90:25-48)
    salt_node = np.zeros((L, P), order="F")
    temp_node = np.full((L, P), 20.0, order="F")
    if case.is_salt and case.salt_init is not None:
        salt_node = np.asarray(case.salt_init, dtype=np.float64).T.copy(order="F")
    elif case.baroclinic_flag == 1 and not case.is_salt:
        salt_node[:] = case.ref_salt
    if case.is_temp and case.temp_init is not None:
        temp_node = np.asarray(case.temp_init, dtype=np.float64).T.copy(order="F")
    elif case.baroclinic_flag == 1 and not case.is_temp:
        temp_node[:] = case.ref_temp
    nf = mesh.view2("nodenum_at_face")

    def face_mean(a):
        return np.asfortranarray((a[:, nf[0] - 1] + a[:, nf[1] - 1]) * 0.5)

    def cell_mean(a):
        acc = np.zeros((L, E), order="F")
        for l in range(4):
            m_ = l < N
            acc = np.where(m_[None, :], acc + a[:, np.maximum(nc[l], 1) - 1], acc)
        return np.asfortranarray(acc / N[None, :])
    if case.is_salt and case.salt_init is not None:
        salt_face, salt_cell = face_mean(salt_node), cell_mean(salt_node)
    else:
        salt_face = np.full((L, J), salt_node[0, 0] if P else 0.0, order="F")
        salt_cell = np.full((L, E), salt_node[0, 0] if P else 0.0, order="F")
    if case.is_temp and case.temp_init is not None:
        temp_face, temp_cell = face_mean(temp_node), cell_mean(temp_node)
    else:
        temp_face = np.full((L, J), temp_node[0, 0] if P else 20.0, order="F")
        temp_cell = np.full((L, E), temp_node[0, 0] if P else 20.0, order="F")
    for nm, v in (("salt_node", salt_node), ("salt_face", salt_face), ("salt_cell", salt_cell),
                  ("salt_cell_new", salt_cell), ("temp_node", temp_node), ("temp_face", temp_face),
                  ("temp_cell", temp_cell), ("temp_cell_new", temp_cell)):
        s[nm] = v
    return s


def hot_start(case: Case, p: "abi.Params", mesh: "abi.Arrays") -> "abi.Arrays":
    """Initial state after prepare_gom with restart_in = 1 (EOS excluded).

    prepare_gom.f90:80-92 builds the layers from the surface as it stands before the
    restart is read (zero), and setup_hot_start.f90:14-62 then overwrites only the restart
    fields. The initial tracer files are not read (read_main_inp.f90:316-321), so the
    tracers the restart does not hold keep their allocation values (salt 0, temperature
    20; read_main_inp.f90:290-303 for a tracer that is not transported)."""
    from dataclasses import replace
    from .gomout import RESTART_FIELDS
    if case.hot_start is None:
        raise ValueError("restart_in = 1 needs the restart records (Case.hot_start)")
    s = cold_start(replace(case, salt_init=None, temp_init=None), p, mesh)
    for f in RESTART_FIELDS:
        s[f] = case.hot_start[f]
    return s
