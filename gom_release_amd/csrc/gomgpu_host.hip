// gomgpu_host.hip - the C ABI of include/gomgpu.h: context, uploads/downloads in the
// reference's Fortran layout, and the stage sequence of one time step
// (GOM_Rv1.1.0/source/main.f90:95-272).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "gomgpu_dev.h"

// kernels (k_hydro.hip, k_transport.hip, k_pcg.hip)
__global__ void k_friction(DevMesh, DevState, DevDims, int, double, double);
__global__ void k_elm_substeps(DevMesh, DevState, DevDims, double, double, int, int);
__global__ void k_elm(DevMesh, DevState, DevDims, double, double, double, int *, int32_t *, unsigned long long *);
__global__ void k_elm_copy(DevState, long);
__global__ void k_elm_tracer(DevMesh, DevState, DevForcing, DevWork, DevDims, double, double, double, int, int *);
__global__ void k_elm_column(DevMesh, DevState, DevWork, DevDims, double, int, int, int *);
__global__ void k_elm_node(DevMesh, DevState, DevDims);
__global__ void k_elm_face(DevMesh, DevState, DevDims);
__global__ void k_cell_uv(DevMesh, DevState, DevWork, DevDims);
__global__ void k_mom_rhs(DevMesh, DevState, DevForcing, DevWork, DevDims, MomParams);
__global__ void k_mom_solve(DevMesh, DevState, DevForcing, DevWork, DevDims, MomParams, int *);
__global__ void k_river_velocity(DevMesh, DevState, DevForcing, DevDims, double *);
__global__ void k_hurricane_nodes(DevForcing, int, double);
__global__ void k_holland_next(DevMesh, DevForcing, gomgpu_holland, int);
__global__ void k_wind_faces(DevMesh, DevForcing, int, double, int);
__global__ void k_fs_face(DevMesh, DevState, DevWork, DevDims, double);
__global__ void k_fs_cell(DevMesh, DevState, DevForcing, DevWork, DevDims, double, double, int *);
__global__ void k_eta_node(DevMesh, DevState, DevDims);
__global__ void k_eta_node_dry(DevMesh, DevState, DevDims);
__global__ void k_velocity_face(DevMesh, DevState, DevForcing, DevDims, double);
__global__ void k_velocity_w(DevMesh, DevState, DevDims);
__global__ void k_layers_cell(DevMesh, DevState, DevDims, double, double, int *);
__global__ void k_layers_face(DevMesh, DevState, DevDims, double, double, int *);
__global__ void k_layers_node(DevMesh, DevState, DevDims, double, double, int *);
__global__ void k_layers_check_cell(DevMesh, DevState, DevDims, double, int *);
__global__ void k_layers_check_fn(DevMesh, DevState, DevDims, double, int *);
__global__ void k_face_levels(DevMesh, DevState, DevDims);
__global__ void k_node_vel(DevMesh, DevState, DevDims);
// transport launchers (k_transport.hip): f32 selects the FP32-storage instantiation
void tvd_launch_init(bool, long, int, hipStream_t, DevMesh, DevState, DevForcing, DevWork, DevDims, TvdArgs);
void tvd_launch_bnd(bool, long, int, hipStream_t, DevMesh, DevState, DevForcing, DevWork, DevDims, TvdArgs);
void tvd_launch_coef_face(bool, long, int, hipStream_t, DevMesh, DevState, DevWork, DevDims, TvdArgs);
void tvd_launch_coef_cell(bool, long, int, hipStream_t, DevMesh, DevState, DevWork, DevDims, TvdArgs);
void tvd_launch_lim_cell(bool, long, int, hipStream_t, DevMesh, DevState, DevWork, DevDims, int, int);
void tvd_launch_lim_face(bool, long, int, hipStream_t, DevMesh, DevState, DevWork, DevDims, int, int);
void tvd_launch_rhs(bool, long, int, hipStream_t, DevMesh, DevState, DevWork, DevDims, TvdArgs);
void tvd_launch_solve(bool, long, int, hipStream_t, DevMesh, DevState, DevWork, DevDims, TvdArgs, double *, int, int,
                      int *);
void tvd_launch_node(bool, long, int, hipStream_t, DevMesh, DevState, DevWork, DevDims);
void tvd_launch_face_out(bool, long, int, hipStream_t, DevMesh, DevState, DevWork, DevDims);
void tvd_launch_cellnode_out(bool, long, int, hipStream_t, DevState, DevWork, DevDims);
__global__ void k_eos_nodeface(DevMesh, DevState, DevDims, double, int, int, int *);
__global__ void k_eos_cell(DevMesh, DevState, DevDims, int *);
__global__ void k_bars(DevMesh, DevState, DevDims);
__global__ void k_turb_ri(DevMesh, DevState, DevWork, DevDims, TurbParams);
__global__ void __launch_bounds__(256) k_turb_column(DevMesh, DevState, DevForcing, DevWork, DevDims, TurbParams, int *);
__global__ void __launch_bounds__(256) k_turb_levels(DevMesh, DevState, DevForcing, DevWork, DevDims, TurbParams, int *);
__global__ void k_turb_kv(DevMesh, DevState, DevWork, DevDims);
__global__ void k_budget(DevMesh, DevState, DevDims, double *);
__global__ void k_copy(double *, const double *, int);

// Segments of the packed forcing upload: segment i goes to dst[i] from offset off[i].
struct ScatterSegs {
  double *dst[GOM_MAX_FORCING + 1];
  long off[GOM_MAX_FORCING + 1], n[GOM_MAX_FORCING + 1];
  int count;
};

#define CNT_F(name, ext) +1
static_assert(0 GOMGPU_FORCING_REAL(CNT_F) <= GOM_MAX_FORCING, "GOM_MAX_FORCING below the forcing array count");
#undef CNT_F

// One block row (blockIdx.y) per segment, grid-strided over its values.
__global__ void k_scatter(ScatterSegs s, const double *src) {
  const int g = blockIdx.y;
  const long n = s.n[g], o = s.off[g];
  double *dst = s.dst[g];
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    dst[i] = src[o + i];
}
__global__ void k_update_dz(double *, const double *, double *, const double *, const int32_t *, const int32_t *, int, int);
__global__ void k_fill(double *, double, long);
extern "C" hipError_t gom_launch_pcg(const PcgArgs *args, int max_blocks, int cooperative, hipStream_t st);
extern "C" int gom_pcg_max_blocks(int device);
extern "C" hipError_t gom_launch_pcg_reg(const PcgArgs *args, int nb, int bs, int cooperative, hipStream_t st);
extern "C" int gom_pcg_reg_capacity(int device, int n, int nb, int bs);
extern "C" hipError_t gom_launch_pcg_cg1(const PcgArgs *args, int nb, int bs, int slots, int cooperative,
                                         hipStream_t st);
extern "C" int gom_pcg_cg1_capacity(int device, int n, int nb, int bs, int slots);
extern "C" int gom_pcg_cg1m_blocks(long n);
extern "C" hipError_t gom_launch_pcg_cg1m(const PcgArgs *args, int nb, int cooperative, hipStream_t st);
extern "C" int gom_pcg_cg1m_capacity(int device);
extern "C" hipError_t gom_launch_pcg_mixed(const PcgArgs *a64, const PcgArgs32 *m32, int cooperative,
                                           hipStream_t st);
extern "C" int gom_pcg_cg1m32_capacity(int device);
extern "C" hipError_t gom_launch_boundary(const DevBnd *b, const DevForcing *f, const DevDims *d, double dt, int Nt,
                                          int elm, hipStream_t st);

enum { KT_PCG = 0, KT_ELM = 1, KT_MOMENTUM = 2, KT_TVD_COLUMN = 3 };
static const char *kKernelNames[GOMGPU_NUM_KTIMERS] = {"k_pcg", "k_elm", "k_momentum", "k_tvd_column"};
static const char *kStageNames[GOMGPU_NUM_STAGES] = {"friction", "elm",       "momentum", "freesurf",
                                                      "velocity", "layers",    "nodevel",  "transport",
                                                      "eos",      "update",    "turbulence", "wind",
                                                      "boundary", "bars"};
// Order of the stages in one step (main.f90:95-272, the closure before the transport; the
// boundary forcing and the wind stage of main.f90:84-161 first, each a no-op unless its
// tables or hurricane fields are set).
static const int kStepOrder[] = {
    GOMGPU_STAGE_BOUNDARY,       GOMGPU_STAGE_WIND,
    GOMGPU_STAGE_FRICTION, GOMGPU_STAGE_ELM,        GOMGPU_STAGE_MOMENTUM,  GOMGPU_STAGE_FREESURF,
    GOMGPU_STAGE_VELOCITY, GOMGPU_STAGE_LAYERS,     GOMGPU_STAGE_NODEVEL,   GOMGPU_STAGE_TURBULENCE,
    GOMGPU_STAGE_TRANSPORT, GOMGPU_STAGE_EOS,       GOMGPU_STAGE_UPDATE};
static const int kNumStep = (int)(sizeof(kStepOrder) / sizeof(kStepOrder[0]));

struct gomgpu_ctx {
  int device = 0;
  hipStream_t st = nullptr;
  gomgpu_params p;
  DevDims d;
  DevMesh m;
  DevState s;
  DevForcing f;
  DevWork w;
  double elapsed = 0.0, spin_tide = 1.0, spin_bc = 1.0;
  std::vector<void *> allocs;
  size_t bytes = 0;  // device memory held by the context
  // threads per block of the tunable launches (GOMGPU_LAUNCH_*); none changes results
  int lb[GOMGPU_NUM_LAUNCH] = {128, 256, 64, 256, 128, 256, 256, 256, 128, 256};
  int pcg_blocks = 0;      // co-resident capacity of the PCG kernel
  int pcg_cap = 0;         // tuned cap of the PCG grid (0 = capacity)
  int pcg_barrier = 1;     // 0 cooperative-groups sync, 1 custom barrier
  int pcg_coop = 1;        // launch through hipLaunchCooperativeKernel
  int pcg_mode = 3;        // 0 strided k_pcg, 1 register-resident k_pcg_reg, 2 single-barrier k_pcg_cg1,
                           // 3 k_pcg_cg1 with the slot exchange instead of the grid barrier (and
                           // k_pcg_cg1m above its register capacity), 4 always k_pcg_cg1m
  int pcg_bs = 512;        // threads per block of k_pcg_reg / k_pcg_cg1
  int pcg_nb = 0;          // blocks of k_pcg_reg (0 = rows / (4 * bs))
  int pcg_rows = 0;        // rows per thread of the register-resident solve (0 = automatic)
  int ncu = 0;
  int last_iters = 0;
  std::string err;
  bool timing = false;
  hipEvent_t ev[GOMGPU_NUM_STAGES + 1] = {};
  double stage_ms[GOMGPU_NUM_STAGES];
  // per-kernel timers of the hot kernels (event pairs on the context stream)
  hipEvent_t kev[GOMGPU_NUM_KTIMERS][2] = {};
  double kernel_ms[GOMGPU_NUM_KTIMERS];
  int kernel_calls[GOMGPU_NUM_KTIMERS];
  int *h_err = nullptr;  // pinned
  bool pcg_ran = false;  // a free-surface solve ran since the last sync
  std::vector<gomgpu_batch *> batches;  // batches holding this context (a captured batch step
                                        // holds its pointers)
  unsigned stale = 0;  // SWAP_* pairs whose new array was exchanged and not yet rewritten
  // The layer thicknesses are copied new -> old over the levels the last LAYERS stage wrote,
  // which is exact when new == old held everywhere before it: not after set_state, nor
  // when LAYERS ran other than once since the last update (then every entry is copied).
  bool dz_full = true;
  int layers_since_update = 0;
  double *h_stage = nullptr;  // pinned staging of the forcing arrays
  double *d_stage = nullptr;  // its device image: one copy, then k_scatter to the fields
  size_t stage_len = 0;
  hipEvent_t fev = nullptr;   // completion of the last forcing upload
  // boundary forcing on the device (gomgpu_set_boundary): its tables and the sub-cycle
  // extent of con_at_ob / con_at_Qb
  bool bnd_on = false;
  DevBnd bnd{};
  int bnd_nt = 1;
  // a Holland cyclone generated on the device (gomgpu_set_holland)
  bool holland_on = false;
  gomgpu_holland holland{};
  // asynchronous state snapshots (gomgpu_snapshot): a device copy of the fields taken on the
  // step's stream, then copied to pinned host memory on a stream of its own
  struct SnapItem {
    void *dst;
    size_t off;  // bytes into the snapshot buffers
    long ld, n;
    bool is_int;
  };
  std::mutex snap_mu;
  hipStream_t snap_st = nullptr;
  hipEvent_t snap_ev0 = nullptr, snap_ev = nullptr;
  char *snap_dev = nullptr, *snap_host = nullptr;
  size_t snap_cap = 0;
  bool snap_busy = false;
  std::vector<SnapItem> snap_items;
  int32_t *dep_buf = nullptr; // departure cells and levels of the momentum ELM (allocated on
                              // the first gomgpu_elm_record; w.elm_dep points here while on)
};

static int seterr(gomgpu_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, call)                                                                     \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return seterr(ctx, GOMGPU_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));    \
  } while (0)

// ------------------------------------------------------------------ shapes
struct Shape {
  long ld, n;  // Fortran leading dimension (slots/layers) and entity count
};

static Shape shape_of(const char *code, const gomgpu_params &p) {
  long P = p.maxnod, E = p.maxele, J = p.maxface, L = p.maxlayer, NB = p.max_no_neighbor_node;
  long OB = p.num_ob_cell > 0 ? p.num_ob_cell : 1, QB = p.num_Qb_cell > 0 ? p.num_Qb_cell : 1;
  long NT = p.maxtran2 > 0 ? p.maxtran2 : 1, Nt = p.trans_sub_iter > 0 ? p.trans_sub_iter : 1;
  std::string c(code);
  if (c == "P") return {1, P};
  if (c == "E") return {1, E};
  if (c == "J") return {1, J};
  if (c == "L") return {1, L};
  if (c == "L1") return {1, L + 1};
  if (c == "4E") return {4, E};
  if (c == "2J") return {2, J};
  if (c == "NBP") return {NB, P};
  if (c == "NB1P") return {NB + 1, P};
  if (c == "OB") return {1, OB};
  if (c == "QB") return {1, QB};
  if (c == "LJ") return {L, J};
  if (c == "L1J") return {L + 1, J};
  if (c == "LE") return {L, E};
  if (c == "L1E") return {L + 1, E};
  if (c == "LP") return {L, P};
  if (c == "L1P") return {L + 1, P};
  if (c == "NTOB") return {1, Nt * NT * OB};
  if (c == "NTQB") return {1, Nt * NT * QB};
  if (c == "HP") return {1, p.hurricane_flag == 1 ? P : 1};
  if (c == "HJ") return {1, p.hurricane_flag == 1 ? J : 1};
  return {0, 0};
}

template <typename T>
static void to_device_layout(const T *src, T *dst, Shape sh) {
  // Fortran (ld, n) column-major -> [slot][entity]
  for (long e = 0; e < sh.n; e++)
    for (long l = 0; l < sh.ld; l++) dst[l * sh.n + e] = src[l + e * sh.ld];
}

template <typename T>
static void from_device_layout(const T *src, T *dst, Shape sh) {
  for (long e = 0; e < sh.n; e++)
    for (long l = 0; l < sh.ld; l++) dst[l + e * sh.ld] = src[l * sh.n + e];
}

template <typename T>
static int dalloc(gomgpu_ctx *c, T **ptr, size_t count) {
  void *p = nullptr;
  if (count == 0) count = 1;
  if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) return seterr(c, GOMGPU_E_NOMEM, "hipMalloc failed");
  // zeroed on the context's own stream: the stream is non-blocking, so a fill on the null
  // stream would not be ordered before the uploads and kernels that follow on it
  if (hipMemsetAsync(p, 0, count * sizeof(T), c->st) != hipSuccess) return seterr(c, GOMGPU_E_HIP, "hipMemsetAsync failed");
  c->allocs.push_back(p);
  c->bytes += count * sizeof(T);
  *ptr = (T *)p;
  return 0;
}

template <typename T>
static int upload(gomgpu_ctx *c, T *dptr, const T *host, Shape sh) {
  std::vector<T> tmp((size_t)(sh.ld * sh.n));
  to_device_layout(host, tmp.data(), sh);
  HIPCHK(c, hipMemcpyAsync(dptr, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return 0;
}

template <typename T>
static int download(gomgpu_ctx *c, T *host, const T *dptr, Shape sh) {
  std::vector<T> tmp((size_t)(sh.ld * sh.n));
  HIPCHK(c, hipMemcpyAsync(tmp.data(), dptr, tmp.size() * sizeof(T), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  from_device_layout(tmp.data(), host, sh);
  return 0;
}

// ------------------------------------------------------------------ create
static int validate(const gomgpu_params *p, std::string &why) {
  if (p->maxnod <= 0 || p->maxele <= 0 || p->maxface <= 0) return why = "empty mesh", 1;
  if (p->maxlayer < 1 || p->maxlayer > GOM_LMAX) return why = "maxlayer must be in [1, 64]", 1;
  if (p->max_no_neighbor_node < 1) return why = "max_no_neighbor_node < 1", 1;
  if (p->transport_flag == 1) {
    if (p->maxtran2 < 1 || p->maxtran2 > 2) return why = "maxtran2 must be 1 or 2", 1;
    if (p->tran_id1 != 1)
      return why = "the first transported tracer must be salinity (the reference copies the "
                   "temperature result from tracer slot 2)", 1;
    if (p->maxtran2 == 2 && p->tran_id2 != 2) return why = "tracer 2 must be temperature", 1;
    if (p->heat_option == 2) return why = "heat_option 2 is not supported", 1;
    if (p->transport_solver < 1 || p->transport_solver > 2) return why = "transport_solver must be 1 or 2", 1;
    if (p->transport_solver == 2 && p->maxtran2 != 2)
      return why = "the ELM transport solver (transport_solver 2) carries salinity and temperature: maxtran2 "
                   "must be 2", 1;
  }
  if (p->baroclinic_flag == 1 && p->ana_density != 0 && p->ana_density != 1)
    return why = "ana_density must be 0 or 1", 1;
  if (p->turbulence_flag < 0 || p->turbulence_flag > 3) return why = "turbulence_flag must be 0..3", 1;
  if (p->turbulence_flag == 3) {
    if (p->turbulence_model < 0 || p->turbulence_model > 4) return why = "turbulence_model must be 0..4", 1;
    if (p->stability_function < 0 || p->stability_function > 1) return why = "stability_function must be 0 or 1", 1;
    if (!(p->q2min > 0.0) || !(p->xlmin2 > 0.0) || !(p->eps_min > 0.0) || p->rnub == 0.0)
      return why = "closure needs q2min > 0, xlmin2 > 0, eps_min > 0 and rnub != 0", 1;
  }
  if (p->hurricane_flag != 0 && p->hurricane_flag != 1) return why = "hurricane_flag must be 0 or 1", 1;
  if (p->hurricane_flag == 1 && p->hurricane_read_interval < 1)
    return why = "hurricane_read_interval (hurricane_dt / dt) must be at least 1", 1;
  if (p->precision != GOMGPU_FP64 && p->precision != GOMGPU_MIXED) return why = "bad precision", 1;
  return 0;
}

// Why the last gomgpu_create on this thread failed (gomgpu_last_error(NULL)).
static thread_local std::string g_create_err;

extern "C" int gomgpu_create(const gomgpu_mesh *mesh, const gomgpu_params *params, gomgpu_ctx **out) {
  if (!mesh || !params || !out) {
    g_create_err = "gomgpu_create: NULL mesh, params or out";
    return GOMGPU_E_ARG;
  }
  *out = nullptr;
  g_create_err.clear();
  std::string why;
  if (validate(params, why)) {
    g_create_err = "gomgpu_create: " + why;
    fprintf(stderr, "%s\n", g_create_err.c_str());
    return GOMGPU_E_ARG;
  }
  gomgpu_ctx *c = new gomgpu_ctx();
  c->p = *params;
  const gomgpu_params &p = c->p;
  int rc;
  {
    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
    if (e != hipSuccess) {
      g_create_err = std::string("gomgpu_create: no usable HIP device: ") + hipGetErrorString(e);
      fprintf(stderr, "%s\n", g_create_err.c_str());
      delete c;
      return GOMGPU_E_HIP;
    }
  }
  c->d.P = p.maxnod;
  c->d.E = p.maxele;
  c->d.J = p.maxface;
  c->d.L = p.maxlayer;
  c->d.NB = p.max_no_neighbor_node;
  c->d.NT = p.transport_flag == 1 ? p.maxtran2 : 0;
  c->d.NOB = p.num_ob_cell;
  c->d.NQB = p.num_Qb_cell;
  c->d.tid[0] = p.tran_id1;
  c->d.tid[1] = p.tran_id2;
  // mesh
#define UP_MR(name, ext)                                                             \
  {                                                                                  \
    Shape sh = shape_of(#ext, p);                                                    \
    double *dp;                                                                      \
    if ((rc = dalloc(c, &dp, sh.ld * sh.n))) goto fail;                              \
    if (mesh->name && (rc = upload(c, dp, mesh->name, sh))) goto fail;               \
    c->m.name = dp;                                                                  \
  }
#define UP_MI(name, ext)                                                             \
  {                                                                                  \
    Shape sh = shape_of(#ext, p);                                                    \
    int32_t *dp;                                                                     \
    if ((rc = dalloc(c, &dp, sh.ld * sh.n))) goto fail;                              \
    if (mesh->name && (rc = upload(c, dp, mesh->name, sh))) goto fail;               \
    c->m.name = dp;                                                                  \
  }
  GOMGPU_MESH_REAL(UP_MR)
  GOMGPU_MESH_INT(UP_MI)
#undef UP_MR
#undef UP_MI
  {
    // Manning drag of the still-water face depth (calculate_bottom_friction.f90:27-37),
    // evaluated with the host libm pow as the reference does.
    std::vector<double> cdb(p.maxface, 0.0);
    if (p.bf_flag == 1) {
      const double expo = 1.0 / 6.0;
      for (int j = 0; j < p.maxface; j++) {
        double hr = 1.0 > mesh->h_face[j] ? 1.0 : mesh->h_face[j];
        double chezy = pow(hr, expo) / mesh->manning[j];
        cdb[j] = p.gravity / (chezy * chezy);
      }
    }
    double *dp;
    if ((rc = dalloc(c, &dp, p.maxface))) goto fail;
    if ((rc = upload(c, dp, cdb.data(), Shape{1, p.maxface}))) goto fail;
    c->m.Cdb_static = dp;
  }
  {
    // node stencil (calculate_velocity_at_node.f90:67-138): the faces and cells the node
    // loop of the reference visits, in its order, with the level-independent factors
    const long P = p.maxnod, NB = p.max_no_neighbor_node;
    if (!mesh->adj_cells_at_node || !mesh->adj_cellnum_at_node || !mesh->node_count_each_element ||
        !mesh->tri_or_quad || !mesh->facenum_at_cell || !mesh->adj_cellnum_at_face ||
        !mesh->bottom_layer_at_face || !mesh->bottom_layer_at_element || !mesh->face_length || !mesh->area) {
      rc = seterr(c, GOMGPU_E_ARG, "mesh connectivity missing");
      goto fail;
    }
    std::vector<int32_t> face(4 * NB * P, 0), cell(2 * NB * P, 0);
    std::vector<double> Lf(4 * NB * P, 0.0), area(NB * P, 0.0), weight(P, 0.0);
    for (size_t q = 0; q < Lf.size(); q += 4) Lf[q] = Lf[q + 1] = 1.0;  // lengths of unused slots
    for (long i = 0; i < P; i++) {
      int nadj = mesh->adj_cells_at_node[i];
      double wsum = 0.0;
      for (int jj = 1; jj <= nadj && jj <= NB; jj++) {
        int isn = mesh->adj_cellnum_at_node[(jj - 1) + i * NB] - 1;
        int nc = mesh->node_count_each_element[(jj - 1) + i * NB];
        int N = mesh->tri_or_quad[isn];
        for (int l = N - 2; l <= N - 1; l++) {
          int v = nc + l;
          int sl = v > N ? v - N : v;  // start_end_node (set_geometry_1.f90:78-97)
          int ifn = mesh->facenum_at_cell[(sl - 1) + (long)isn * 4] - 1;
          double w = mesh->adj_cellnum_at_face[1 + (long)ifn * 2] == 0 ? 2.0 : 1.0;
          // one record per (adjacent cell, node): (face 0, face 1, kb 0, kb 1) and
          // (L 0, L 1, w 0, w 1), the two faces of the cell that touch the node
          const int h = l - (N - 2);
          long o = 4 * ((long)(jj - 1) * P + i);
          face[o + h] = ifn;
          face[o + 2 + h] = mesh->bottom_layer_at_face[ifn];
          Lf[o + h] = mesh->face_length[ifn];
          Lf[o + 2 + h] = w;
          wsum = wsum + w / mesh->face_length[ifn];
        }
        long oc = (long)(jj - 1) * P + i;
        cell[2 * oc] = isn;
        cell[2 * oc + 1] = mesh->bottom_layer_at_element[isn];
        area[oc] = mesh->area[isn];
      }
      weight[i] = wsum;
    }
    int32_t *ip;
    double *dp;
#define NS_UP(vec, field, T, ptr)                                                     \
    if ((rc = dalloc(c, &ptr, vec.size())) || (rc = upload(c, ptr, vec.data(), Shape{1, (long)vec.size()}))) goto fail; \
    c->m.field = ptr;
    NS_UP(face, ns_face, int32_t, ip)
    NS_UP(cell, ns_cell, int32_t, ip)
    NS_UP(Lf, ns_Lf, double, dp)
    NS_UP(area, ns_area, double, dp)
    NS_UP(weight, ns_weight, double, dp)
#undef NS_UP
  }
  {
    // cell records of the ELM walk: node coordinates and ids in slot order, 4 per cell
    const long E = p.maxele;
    if (!mesh->nodenum_at_cell || !mesh->x_node || !mesh->y_node) {
      rc = seterr(c, GOMGPU_E_ARG, "mesh connectivity missing");
      goto fail;
    }
    if (!mesh->adj_cellnum_at_cell) {
      rc = seterr(c, GOMGPU_E_ARG, "mesh connectivity missing");
      goto fail;
    }
    std::vector<double> cx(4 * E, 0.0), cy(4 * E, 0.0);
    std::vector<int32_t> cn(4 * E, 0), ca(4 * E, 0);
    for (long e = 0; e < E; e++) {
      int N = mesh->tri_or_quad[e];
      for (int l = 0; l < N && l < 4; l++) {
        int nd = mesh->nodenum_at_cell[l + e * 4];
        cn[4 * e + l] = nd;
        ca[4 * e + l] = mesh->adj_cellnum_at_cell[l + e * 4];
        cx[4 * e + l] = mesh->x_node[nd - 1];
        cy[4 * e + l] = mesh->y_node[nd - 1];
      }
    }
    int32_t *ip;
    double *dp;
    if ((rc = dalloc(c, &dp, cx.size())) || (rc = upload(c, dp, cx.data(), Shape{1, (long)cx.size()}))) goto fail;
    c->m.cg_x = dp;
    if ((rc = dalloc(c, &dp, cy.size())) || (rc = upload(c, dp, cy.data(), Shape{1, (long)cy.size()}))) goto fail;
    c->m.cg_y = dp;
    if ((rc = dalloc(c, &ip, cn.size())) || (rc = upload(c, ip, cn.data(), Shape{1, (long)cn.size()}))) goto fail;
    c->m.cg_node = ip;
    if ((rc = dalloc(c, &ip, ca.size())) || (rc = upload(c, ip, ca.data(), Shape{1, (long)ca.size()}))) goto fail;
    c->m.cg_adj = ip;
    // slot_face: each cell slot's face and outflow sign where the transport counts its flux
    if (!mesh->facenum_at_cell || !mesh->sign_in_outflow || !mesh->boundary_type_of_face || !mesh->isflowside3) {
      rc = seterr(c, GOMGPU_E_ARG, "mesh connectivity missing");
      goto fail;
    }
    std::vector<int32_t> sf(4 * E, 0);
    for (long e = 0; e < E; e++) {
      int N = mesh->tri_or_quad[e];
      for (int l = 0; l < N && l < 4; l++) {
        int j = mesh->facenum_at_cell[l + e * 4] - 1;
        int sg = mesh->sign_in_outflow[l + e * 4];
        if (j < 0 || j >= p.maxface || (sg != 1 && sg != -1)) {
          rc = seterr(c, GOMGPU_E_ARG, "sign_in_outflow must be +1 or -1 on every face of every cell");
          goto fail;
        }
        const bool active = mesh->boundary_type_of_face[j] != -1 || mesh->isflowside3[j] > 0;
        sf[4 * e + l] = active ? (int32_t)(((uint32_t)(j + 1) << 1) | (sg > 0 ? 1u : 0u)) : 0;
      }
    }
    if ((rc = dalloc(c, &ip, sf.size())) || (rc = upload(c, ip, sf.data(), Shape{1, (long)sf.size()}))) goto fail;
    c->m.slot_face = ip;
    // face_adj: per face its two cells (1-based, 0 = none), the first shifted left by one
    // with bit 0 set where the transport counts the face's flux
    std::vector<int32_t> fa(2 * (size_t)p.maxface);
    for (long j = 0; j < p.maxface; j++) {
      const int32_t c1 = mesh->adj_cellnum_at_face[2 * j], c2 = mesh->adj_cellnum_at_face[2 * j + 1];
      const bool active = mesh->boundary_type_of_face[j] != -1 || mesh->isflowside3[j] > 0;
      fa[2 * j] = (int32_t)(((uint32_t)c1 << 1) | (active ? 1u : 0u));
      fa[2 * j + 1] = c2;
    }
    if ((rc = dalloc(c, &ip, fa.size())) || (rc = upload(c, ip, fa.data(), Shape{1, (long)fa.size()}))) goto fail;
    c->m.face_adj = ip;
  }
  // state
#define AL_SR(name, ext)                                                             \
  {                                                                                  \
    Shape sh = shape_of(#ext, p);                                                    \
    if ((rc = dalloc(c, &c->s.name, sh.ld * sh.n))) goto fail;                       \
  }
  GOMGPU_STATE_REAL(AL_SR)
  GOMGPU_STATE_INT(AL_SR)
#undef AL_SR
  // forcing
#define AL_FR(name, ext)                                                             \
  {                                                                                  \
    Shape sh = shape_of(#ext, p);                                                    \
    double *dp;                                                                      \
    if ((rc = dalloc(c, &dp, sh.ld * sh.n))) goto fail;                              \
    c->f.name = dp;                                                                  \
  }
  GOMGPU_FORCING_REAL(AL_FR)
#undef AL_FR
  {
    double *dp;
    const double init[GOM_NSCAL] = {1.0, 1.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0};
    if ((rc = dalloc(c, &dp, GOM_NSCAL))) goto fail;
    if (hipMemcpyAsync(dp, init, sizeof init, hipMemcpyHostToDevice, c->st) != hipSuccess ||
        hipStreamSynchronize(c->st) != hipSuccess) {
      rc = seterr(c, GOMGPU_E_HIP, "hipMemcpy failed");
      goto fail;
    }
    c->f.scal = dp;
  }
  {
    long E = p.maxele, J = p.maxface, P = p.maxnod, L = p.maxlayer;
    long NT = c->d.NT > 0 ? c->d.NT : 1, OB = p.num_ob_cell > 0 ? p.num_ob_cell : 1,
         QB = p.num_Qb_cell > 0 ? p.num_Qb_cell : 1;
    c->pcg_blocks = gom_pcg_max_blocks(c->device);
    hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, c->device);
    if (const char *e = getenv("GOMGPU_PCG_BLOCKS")) c->pcg_cap = atoi(e);
    if (const char *e = getenv("GOMGPU_PCG_BARRIER")) c->pcg_barrier = atoi(e);
    if (const char *e = getenv("GOMGPU_PCG_COOPERATIVE")) c->pcg_coop = atoi(e);
    if (const char *e = getenv("GOMGPU_PCG_MODE")) c->pcg_mode = atoi(e);
    if (const char *e = getenv("GOMGPU_PCG_BS")) c->pcg_bs = atoi(e);
    if (const char *e = getenv("GOMGPU_PCG_NB")) c->pcg_nb = atoi(e);
    if (c->pcg_blocks <= 0) {
      rc = seterr(c, GOMGPU_E_HIP, "cannot size the PCG grid");
      goto fail;
    }
    DevWork &w = c->w;
    if ((rc = dalloc(c, &w.fs_diag, E)) || (rc = dalloc(c, &w.fs_off, 4 * E)) || (rc = dalloc(c, &w.fs_rhs, E)) ||
        (rc = dalloc(c, &w.pcg_x, E)) || (rc = dalloc(c, &w.pcg_r, E)) || (rc = dalloc(c, &w.pcg_d0, E)) ||
        (rc = dalloc(c, &w.pcg_d1, E)) || (rc = dalloc(c, &w.pcg_q, E)) || (rc = dalloc(c, &w.pcg_s, E)) ||
        (rc = dalloc(c, &w.pcg_minv, E)) || (rc = dalloc(c, &w.pcg_part, 4 * (long)(c->pcg_blocks > 4096 ? c->pcg_blocks : 4096))) ||
        (rc = dalloc(c, &w.pcg_pub, 6 * E)) || (rc = dalloc(c, &w.pcg_slots, 6 * PCG_MAX_SLOT_BLOCKS)) ||
        (rc = dalloc(c, &w.pcg_info, 3)) || (rc = dalloc(c, &w.pcg_bar, 512)) || (rc = dalloc(c, &w.face_sdz, J)) || (rc = dalloc(c, &w.face_sun, J)) ||
        (rc = dalloc(c, &w.face_sg, J)))
      goto fail;
    if ((rc = dalloc(c, &w.con, NT * (L + 2) * E)) || (rc = dalloc(c, &w.con_new, NT * L * E)) ||
        (rc = dalloc(c, &w.fqd, 2 * L * J)) || (rc = dalloc(c, &w.fps, 2 * L * J)) || (rc = dalloc(c, &w.phij, L * J)) ||
        (rc = dalloc(c, &w.Qi, (L + 1) * E)) || (rc = dalloc(c, &w.di, (L + 1) * E)) ||
        (rc = dalloc(c, &w.phii, (L + 1) * E)) || (rc = dalloc(c, &w.htop, NT * 2 * E)) ||
        (rc = dalloc(c, &w.con2, NT * (L + 2) * E)) ||
        (rc = dalloc(c, &w.PSIi, 2 * (L + 1) * E)) || (rc = dalloc(c, &w.s12, 2 * NT * L * E)) ||
        (rc = dalloc(c, &w.obk, NT * L * OB)) ||
        (rc = dalloc(c, &w.qbk, NT * L * QB)) || (rc = dalloc(c, &w.node_new, NT * L * P)) ||
        (rc = dalloc(c, &w.face_new, NT * L * J)) || (rc = dalloc(c, &w.budget, 2 * ((E + 255) / 256) + 2)) ||
        (rc = dalloc(c, &w.err, 2)) || (rc = dalloc(c, &w.elm_near, 2)) || (rc = dalloc(c, &w.turb_kvf, (L + 1) * J)) ||
        (rc = dalloc(c, &w.tri_gam, L * (E > J ? E : J))) || (rc = dalloc(c, &w.tvd_x, NT * L * E)) ||
        (rc = dalloc(c, &w.cell_u, L * E)) || (rc = dalloc(c, &w.cell_v, L * E)))
      goto fail;
    if (p.turbulence_flag == 3 &&
        ((rc = dalloc(c, &w.tk_shear, L * J)) || (rc = dalloc(c, &w.tk_rz, L * J)) ||
         (rc = dalloc(c, &w.tk_q2tmp, L * J)) || (rc = dalloc(c, &w.tk_x, L * J))))
      goto fail;
  }
  if (hipHostMalloc((void **)&c->h_err, 5 * sizeof(int)) != hipSuccess) {
    rc = seterr(c, GOMGPU_E_NOMEM, "hipHostMalloc failed");
    goto fail;
  }
  for (int i = 0; i <= GOMGPU_NUM_STAGES; i++) hipEventCreate(&c->ev[i]);
  for (int i = 0; i < GOMGPU_NUM_KTIMERS; i++) {
    hipEventCreate(&c->kev[i][0]);
    hipEventCreate(&c->kev[i][1]);
  }
  {
    size_t tot = 0;
#define SZ_F(name, ext)                                                              \
  {                                                                                  \
    Shape sh = shape_of(#ext, p);                                                    \
    tot += (size_t)(sh.ld * sh.n);                                                   \
  }
    GOMGPU_FORCING_REAL(SZ_F)
#undef SZ_F
    tot += GOM_NSCAL;  // the per-step scalars
    c->stage_len = tot;
    if (hipHostMalloc((void **)&c->h_stage, tot * sizeof(double)) != hipSuccess ||
        hipEventCreateWithFlags(&c->fev, hipEventDisableTiming) != hipSuccess ||
        hipEventRecord(c->fev, c->st) != hipSuccess) {
      rc = seterr(c, GOMGPU_E_NOMEM, "cannot allocate the forcing staging buffer");
      goto fail;
    }
    if ((rc = dalloc(c, &c->d_stage, tot))) goto fail;
  }
  if (hipStreamSynchronize(c->st) != hipSuccess) {
    rc = seterr(c, GOMGPU_E_HIP, "initialisation failed");
    goto fail;
  }
  *out = c;
  return 0;
fail:
  g_create_err = "gomgpu_create: " + c->err;
  fprintf(stderr, "%s\n", g_create_err.c_str());
  gomgpu_destroy(c);
  return rc ? rc : GOMGPU_E_HIP;
}

static void batch_forget(gomgpu_batch *b, gomgpu_ctx *c);

extern "C" void gomgpu_destroy(gomgpu_ctx *c) {
  if (!c) return;
  while (!c->batches.empty()) batch_forget(c->batches.back(), c);  // leaves its batches
  if (c->st) hipStreamSynchronize(c->st);
  if (c->snap_st) {
    hipStreamSynchronize(c->snap_st);
    hipStreamDestroy(c->snap_st);
  }
  if (c->snap_ev0) hipEventDestroy(c->snap_ev0);
  if (c->snap_ev) hipEventDestroy(c->snap_ev);
  if (c->snap_dev) hipFree(c->snap_dev);
  if (c->snap_host) hipHostFree(c->snap_host);
  for (void *p : c->allocs) hipFree(p);
  if (c->h_err) hipHostFree(c->h_err);
  if (c->h_stage) hipHostFree(c->h_stage);
  if (c->fev) hipEventDestroy(c->fev);
  for (int i = 0; i <= GOMGPU_NUM_STAGES; i++)
    if (c->ev[i]) hipEventDestroy(c->ev[i]);
  for (int i = 0; i < GOMGPU_NUM_KTIMERS; i++)
    for (int k = 0; k < 2; k++)
      if (c->kev[i][k]) hipEventDestroy(c->kev[i][k]);
  if (c->st) hipStreamDestroy(c->st);
  delete c;
}

// New/old pairs that UPDATE may exchange instead of copying. After an exchange the new
// array holds the values of two steps back until its writer runs (VELOCITY for the
// velocities, TRANSPORT for the transported tracers); until then the pair is stale and means
// new == old, as after the reference's copy. Whatever reads a stale new array first copies
// old into it (materialize), so every entry point sees the reference's values.
enum { SWAP_WN = 1, SWAP_UN = 2, SWAP_VN = 4, SWAP_SALT = 8, SWAP_TEMP = 16, SWAP_ALL = 31 };

static void swap_pair(gomgpu_ctx *c, unsigned bit, double ***old_, double ***new_, size_t *n) {
  const size_t E = c->d.E, J = c->d.J, L = c->d.L;
  switch (bit) {
    case SWAP_WN: *old_ = &c->s.wn_cell; *new_ = &c->s.wn_cell_new; *n = (L + 1) * E; break;
    case SWAP_UN: *old_ = &c->s.un_face; *new_ = &c->s.un_face_new; *n = L * J; break;
    case SWAP_VN: *old_ = &c->s.vn_face; *new_ = &c->s.vn_face_new; *n = L * J; break;
    case SWAP_SALT: *old_ = &c->s.salt_cell; *new_ = &c->s.salt_cell_new; *n = L * E; break;
    default: *old_ = &c->s.temp_cell; *new_ = &c->s.temp_cell_new; *n = L * E; break;
  }
}

static int materialize(gomgpu_ctx *c, unsigned mask) {
  for (unsigned bit = 1; bit <= SWAP_TEMP; bit <<= 1) {
    if (!(c->stale & mask & bit)) continue;
    double **o, **nw;
    size_t n;
    swap_pair(c, bit, &o, &nw, &n);
    k_copy<<<nblk((long)n, 256), 256, 0, c->st>>>(*nw, *o, (int)n);
    HIPCHK(c, hipGetLastError());
    c->stale &= ~bit;
  }
  return 0;
}

extern "C" int gomgpu_set_state(gomgpu_ctx *c, const gomgpu_state *st) {
  if (!c || !st) return GOMGPU_E_ARG;
  const gomgpu_params &p = c->p;
  int rc;
  if ((rc = materialize(c, SWAP_ALL))) return rc;  // fields not given keep new == old
  c->dz_full = true;
#define UP_S(name, ext)                                                              \
  if (st->name) {                                                                    \
    if ((rc = upload(c, c->s.name, st->name, shape_of(#ext, p)))) return rc;         \
  }
  GOMGPU_STATE_REAL(UP_S)
  GOMGPU_STATE_INT(UP_S)
#undef UP_S
  return 0;
}

extern "C" int gomgpu_get_state(gomgpu_ctx *c, gomgpu_state *st) {
  if (!c || !st) return GOMGPU_E_ARG;
  const gomgpu_params &p = c->p;
  int rc;
  if ((rc = materialize(c, SWAP_ALL))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->st));
#define DN_S(name, ext)                                                              \
  if (st->name) {                                                                    \
    if ((rc = download(c, st->name, c->s.name, shape_of(#ext, p)))) return rc;       \
  }
  GOMGPU_STATE_REAL(DN_S)
  GOMGPU_STATE_INT(DN_S)
#undef DN_S
  return 0;
}

// ---- asynchronous snapshots
// The snapshot's device copies go on the context stream (ordered after the work already
// queued, a few microseconds per field at HBM speed); the download runs on the snapshot
// stream, so the steps queued after it do not wait for the bus. One snapshot is in flight per
// context: a new one first completes the previous.
static int snapshot_finish(gomgpu_ctx *c) {
  if (!c->snap_busy) return 0;
  HIPCHK(c, hipEventSynchronize(c->snap_ev));
  for (const auto &it : c->snap_items) {
    Shape sh{it.ld, it.n};
    if (it.is_int) from_device_layout((const int32_t *)(c->snap_host + it.off), (int32_t *)it.dst, sh);
    else from_device_layout((const double *)(c->snap_host + it.off), (double *)it.dst, sh);
  }
  c->snap_items.clear();
  c->snap_busy = false;
  return 0;
}

extern "C" int gomgpu_snapshot(gomgpu_ctx *c, const gomgpu_state *dst) {
  if (!c || !dst) return GOMGPU_E_ARG;
  std::lock_guard<std::mutex> lk(c->snap_mu);
  int rc;
  if ((rc = snapshot_finish(c))) return rc;
  if ((rc = materialize(c, SWAP_ALL))) return rc;
  const gomgpu_params &p = c->p;
  std::vector<gomgpu_ctx::SnapItem> items;
  std::vector<const void *> srcs;
  size_t total = 0;
#define SNAP_F(name, ext, isint)                                                          \
  if (dst->name) {                                                                        \
    Shape sh = shape_of(#ext, p);                                                         \
    items.push_back({(void *)dst->name, total, sh.ld, sh.n, isint});                      \
    srcs.push_back((const void *)c->s.name);                                              \
    total += ((size_t)(sh.ld * sh.n) * (isint ? sizeof(int32_t) : sizeof(double)) + 255) & ~(size_t)255; \
  }
#define SNAP_R(name, ext) SNAP_F(name, ext, false)
#define SNAP_I(name, ext) SNAP_F(name, ext, true)
  GOMGPU_STATE_REAL(SNAP_R)
  GOMGPU_STATE_INT(SNAP_I)
#undef SNAP_R
#undef SNAP_I
#undef SNAP_F
  if (items.empty()) return 0;
  if (!c->snap_st) {
    HIPCHK(c, hipStreamCreateWithFlags(&c->snap_st, hipStreamNonBlocking));
    HIPCHK(c, hipEventCreateWithFlags(&c->snap_ev0, hipEventDisableTiming));
    HIPCHK(c, hipEventCreateWithFlags(&c->snap_ev, hipEventDisableTiming));
  }
  if (total > c->snap_cap) {
    if (c->snap_dev) HIPCHK(c, hipFree(c->snap_dev));
    if (c->snap_host) HIPCHK(c, hipHostFree(c->snap_host));
    c->snap_dev = c->snap_host = nullptr;
    c->snap_cap = 0;
    if (hipMalloc((void **)&c->snap_dev, total) != hipSuccess) return seterr(c, GOMGPU_E_NOMEM, "snapshot hipMalloc failed");
    if (hipHostMalloc((void **)&c->snap_host, total, hipHostMallocDefault) != hipSuccess)
      return seterr(c, GOMGPU_E_NOMEM, "snapshot hipHostMalloc failed");
    c->snap_cap = total;
  }
  for (size_t i = 0; i < items.size(); i++) {
    size_t bytes = (size_t)(items[i].ld * items[i].n) * (items[i].is_int ? sizeof(int32_t) : sizeof(double));
    HIPCHK(c, hipMemcpyAsync(c->snap_dev + items[i].off, srcs[i], bytes, hipMemcpyDeviceToDevice, c->st));
  }
  HIPCHK(c, hipEventRecord(c->snap_ev0, c->st));
  HIPCHK(c, hipStreamWaitEvent(c->snap_st, c->snap_ev0, 0));
  HIPCHK(c, hipMemcpyAsync(c->snap_host, c->snap_dev, total, hipMemcpyDeviceToHost, c->snap_st));
  HIPCHK(c, hipEventRecord(c->snap_ev, c->snap_st));
  c->snap_items = std::move(items);
  c->snap_busy = true;
  return 0;
}

extern "C" int gomgpu_snapshot_wait(gomgpu_ctx *c) {
  if (!c) return GOMGPU_E_ARG;
  std::lock_guard<std::mutex> lk(c->snap_mu);
  return snapshot_finish(c);
}

extern "C" int gomgpu_get_forcing(gomgpu_ctx *c, gomgpu_forcing_out *fo) {
  if (!c || !fo) return GOMGPU_E_ARG;
  const gomgpu_params &p = c->p;
  int rc;
  HIPCHK(c, hipStreamSynchronize(c->st));
#define DN_F(name, ext)                                                                       \
  if (fo->name) {                                                                             \
    if ((rc = download(c, fo->name, (double *)c->f.name, shape_of(#ext, p)))) return rc;      \
  }
  GOMGPU_FORCING_REAL(DN_F)
#undef DN_F
  return 0;
}

// Forcing uploads are asynchronous: the arrays go through one pinned staging buffer; the
// host waits only for the previous forcing copy before reusing it, so it can run one step
// ahead of the device.
extern "C" int gomgpu_set_forcing(gomgpu_ctx *c, const gomgpu_forcing *f) {
  if (!c || !f) return GOMGPU_E_ARG;
  const gomgpu_params &p = c->p;
  c->elapsed = f->elapsed_time;
  c->spin_tide = f->spinup_function_tide;
  c->spin_bc = f->spinup_function_baroclinic;
  HIPCHK(c, hipEventSynchronize(c->fev));
  // the given arrays are packed into the staging buffer, copied in one transfer and spread
  // to their fields by one kernel (one transfer per array costs more stream time)
  size_t off = 0;
  ScatterSegs seg;
  seg.count = 0;
#define UP_F(name, ext)                                                                              \
  if (f->name) {                                                                                     \
    Shape sh = shape_of(#ext, p);                                                                    \
    size_t n = (size_t)(sh.ld * sh.n);                                                               \
    if (off + n > c->stage_len) return seterr(c, GOMGPU_E_ARG, "forcing staging overflow");         \
    to_device_layout(f->name, c->h_stage + off, sh);                                                 \
    seg.dst[seg.count] = (double *)c->f.name;                                                        \
    seg.off[seg.count] = (long)off;                                                                  \
    seg.n[seg.count] = (long)n;                                                                      \
    seg.count++;                                                                                     \
    off += n;                                                                                        \
  }
  GOMGPU_FORCING_REAL(UP_F)
#undef UP_F
  if (off + GOM_NSCAL > c->stage_len) return seterr(c, GOMGPU_E_ARG, "forcing staging overflow");
  if (p.hurricane_flag == 1 && !(f->hurricane_phase == 0.0 || f->hurricane_phase == 1.0 || f->hurricane_phase == 2.0))
    return seterr(c, GOMGPU_E_ARG, "hurricane_phase must be 0, 1 or 2");
  c->h_stage[off] = f->spinup_function_baroclinic;
  c->h_stage[off + 1] = f->spinup_function_tide;
  c->h_stage[off + 2] = f->elapsed_time;
  c->h_stage[off + 3] = f->hurricane_centre_x;
  c->h_stage[off + 4] = f->spinup_function_wind;
  c->h_stage[off + 5] = f->hurricane_phase;
  c->h_stage[off + 6] = f->julian_day;
  c->h_stage[off + 7] = f->hurricane_centre_y;
  seg.dst[seg.count] = (double *)c->f.scal;
  seg.off[seg.count] = (long)off;
  seg.n[seg.count] = GOM_NSCAL;
  seg.count++;
  long nmax = 0;
  for (int i = 0; i < seg.count; i++) nmax = seg.n[i] > nmax ? seg.n[i] : nmax;
  HIPCHK(c, hipMemcpyAsync(c->d_stage, c->h_stage, (off + GOM_NSCAL) * sizeof(double), hipMemcpyHostToDevice,
                           c->st));
  HIPCHK(c, hipEventRecord(c->fev, c->st));
  long bx = (nmax + 255) / 256;
  k_scatter<<<dim3((unsigned)(bx < 1024 ? bx : 1024), (unsigned)seg.count), 256, 0, c->st>>>(seg, c->d_stage);
  HIPCHK(c, hipGetLastError());
  return 0;
}

// Boundary tables (gomgpu_boundary): checked, uploaded once; GOMGPU_STAGE_BOUNDARY then
// computes each step's boundary forcing on the device.
template <typename T>
static int upload_vec(gomgpu_ctx *c, const T **dst, const T *src, size_t n) {
  T *d = nullptr;
  if (int rc = dalloc(c, &d, n)) return rc;
  if (n > 0) HIPCHK(c, hipMemcpyAsync(d, src, n * sizeof(T), hipMemcpyHostToDevice, c->st));
  *dst = d;
  return 0;
}

extern "C" int gomgpu_set_boundary(gomgpu_ctx *c, const gomgpu_boundary *t) {
  if (!c || !t) return GOMGPU_E_ARG;
  const gomgpu_params &p = c->p;
  if (c->bnd_on) return seterr(c, GOMGPU_E_ARG, "gomgpu_set_boundary: the tables are already set");
  if (!c->batches.empty()) return seterr(c, GOMGPU_E_ARG, "gomgpu_set_boundary: the context is in a batch");
  if (t->num_ob != p.num_ob_cell || t->num_qb != p.num_Qb_cell)
    return seterr(c, GOMGPU_E_ARG, "gomgpu_set_boundary: num_ob / num_qb differ from num_ob_cell / num_Qb_cell");
  const int nh = t->num_harmonic_sets, ns = t->num_series, nob = t->num_ob, nqb = t->num_qb;
  if (nh < 0 || ns < 0) return seterr(c, GOMGPU_E_ARG, "gomgpu_set_boundary: negative table size");
  if (nh > 0 && (!t->harmonic_first || !t->tidal_phase_shift || t->harmonic_first[0] != 0))
    return seterr(c, GOMGPU_E_ARG, "gomgpu_set_boundary: harmonic tables missing");
  for (int h = 0; h < nh; h++)
    if (t->harmonic_first[h + 1] < t->harmonic_first[h])
      return seterr(c, GOMGPU_E_ARG, "gomgpu_set_boundary: harmonic_first must not decrease");
  const int ncon = nh > 0 ? t->harmonic_first[nh] : 0;
  if (ncon > 0 && !t->constituents) return seterr(c, GOMGPU_E_ARG, "gomgpu_set_boundary: constituents missing");
  if (ns > 0 && (!t->series_first || t->series_first[0] != 0))
    return seterr(c, GOMGPU_E_ARG, "gomgpu_set_boundary: series tables missing");
  for (int s = 0; s < ns; s++)
    if (t->series_first[s + 1] < t->series_first[s])
      return seterr(c, GOMGPU_E_ARG, "gomgpu_set_boundary: series_first must not decrease");
  const int npt = ns > 0 ? t->series_first[ns] : 0;
  if (npt > 0 && (!t->series_time || !t->series_value))
    return seterr(c, GOMGPU_E_ARG, "gomgpu_set_boundary: series points missing");
  const bool tracers = p.transport_flag == 1 && p.maxtran2 > 0;
  int any_eta_series = 0;
  for (int i = 0; i < nob; i++) {
    const int ty = t->ob_eta_type[i], id = t->ob_eta_id[i];
    if (ty == 1 ? (id < 1 || id > nh) : ty == 2 ? (id < 1 || id > ns) : true)
      return seterr(c, GOMGPU_E_ARG, "gomgpu_set_boundary: ob_eta_type must be 1 (harmonic set) or 2 (series) "
                                     "with an id in range");
    any_eta_series |= ty == 2;
    for (int k = 0; k < 2 && tracers; k++)
      if (t->ob_tracer_series[2 * i + k] < 0 || t->ob_tracer_series[2 * i + k] > ns)
        return seterr(c, GOMGPU_E_ARG, "gomgpu_set_boundary: ob_tracer_series out of range");
  }
  for (int i = 0; i < nqb; i++) {
    if (t->qb_q_series[i] < 0 || t->qb_q_series[i] > ns)
      return seterr(c, GOMGPU_E_ARG, "gomgpu_set_boundary: qb_q_series out of range");
    for (int k = 0; k < 2 && tracers; k++)
      if (t->qb_tracer_series[2 * i + k] < 0 || t->qb_tracer_series[2 * i + k] > ns)
        return seterr(c, GOMGPU_E_ARG, "gomgpu_set_boundary: qb_tracer_series out of range");
  }
  std::vector<int> mono(ns > 0 ? ns : 1, 1);
  for (int s = 0; s < ns; s++)
    for (int k = t->series_first[s] + 1; k < t->series_first[s + 1]; k++)
      if (!(t->series_time[k] >= t->series_time[k - 1])) mono[s] = 0;
  const int Nt = p.transport_solver == 2 ? 1 : (p.trans_sub_iter > 1 ? p.trans_sub_iter : 1);
  DevBnd &b = c->bnd;
  b.nh = nh;
  b.ns = ns;
  b.nob = nob;
  b.nqb = nqb;
  int rc;
  std::vector<int> zeros(2 * (size_t)(nob > nqb ? nob : nqb) + 1, 0);
  if ((rc = upload_vec(c, &b.h_first, t->harmonic_first, nh > 0 ? nh + 1 : 0)) ||
      (rc = upload_vec(c, &b.h_shift, t->tidal_phase_shift, nh)) ||
      (rc = upload_vec(c, &b.h_con, t->constituents, 5 * (size_t)ncon)) ||
      (rc = upload_vec(c, &b.s_first, t->series_first, ns > 0 ? ns + 1 : 0)) ||
      (rc = upload_vec(c, &b.s_mono, mono.data(), ns)) ||
      (rc = upload_vec(c, &b.s_time, t->series_time, npt)) || (rc = upload_vec(c, &b.s_val, t->series_value, npt)) ||
      (rc = upload_vec(c, &b.ob_type, t->ob_eta_type, nob)) || (rc = upload_vec(c, &b.ob_id, t->ob_eta_id, nob)) ||
      (rc = upload_vec(c, &b.ob_ser, tracers ? t->ob_tracer_series : zeros.data(), 2 * (size_t)nob)) ||
      (rc = upload_vec(c, &b.qb_qser, t->qb_q_series, nqb)) || (rc = upload_vec(c, &b.qb_portion, t->qb_portion, nqb)) ||
      (rc = upload_vec(c, &b.qb_ser, tracers ? t->qb_tracer_series : zeros.data(), 2 * (size_t)nqb)))
    return rc;
  if ((rc = dalloc(c, &b.h_old, nh)) || (rc = dalloc(c, &b.h_new, nh)) ||
      (rc = dalloc(c, &b.job_val, (size_t)ns * (2 + Nt))) || (rc = dalloc(c, &b.job_found, (size_t)ns * (2 + Nt))))
    return rc;
  HIPCHK(c, hipStreamSynchronize(c->st));
  b.shift = t->reference_diff_days * 86400.0;
  b.ref_salt = t->ref_salt;
  b.ref_temp = t->ref_temp;
  b.is_salt = t->is_salt;
  b.is_temp = t->is_temp;
  b.any_eta_series = any_eta_series;
  b.tracers = tracers ? 1 : 0;
  c->bnd_nt = Nt;
  c->bnd_on = true;
  return 0;
}

extern "C" int gomgpu_set_holland(gomgpu_ctx *c, const gomgpu_holland *h) {
  if (!c || !h) return GOMGPU_E_ARG;
  if (c->p.hurricane_flag != 1) return seterr(c, GOMGPU_E_ARG, "gomgpu_set_holland: hurricane_flag is not 1");
  if (!c->batches.empty()) return seterr(c, GOMGPU_E_ARG, "gomgpu_set_holland: the context is in a batch");
  if (!(h->r_max > 0.0) || !(h->rho_air > 0.0) || !(h->p_env > h->p_min))
    return seterr(c, GOMGPU_E_ARG, "gomgpu_set_holland: needs r_max > 0, rho_air > 0 and p_env > p_min");
  c->holland = *h;
  c->holland_on = true;
  return 0;
}

// ------------------------------------------------------------------ stages
#define KT_BEGIN(i) \
  if (c->timing) hipEventRecord(c->kev[i][0], st)
#define KT_END(i)                                             \
  if (c->timing) {                                            \
    hipEventRecord(c->kev[i][1], st);                         \
    hipEventSynchronize(c->kev[i][1]);                        \
    float ms_ = 0;                                            \
    hipEventElapsedTime(&ms_, c->kev[i][0], c->kev[i][1]);    \
    c->kernel_ms[i] += ms_;                                   \
    c->kernel_calls[i] += 1;                                  \
  }

// ---- free surface: assembly, PCG, node elevations (solve_free_surface_equation.f90)
static int fs_pre(gomgpu_ctx *c) {
  const gomgpu_params &p = c->p;
  const DevDims &d = c->d;
  const int BS = 256;
  double gdt2th2 = p.gravity * (p.dt * p.dt) * (p.theta * p.theta);
  k_fs_face<<<nblk(d.J, BS), BS, 0, c->st>>>(c->m, c->s, c->w, d, gdt2th2);
  k_fs_cell<<<nblk(d.E, BS), BS, 0, c->st>>>(c->m, c->s, c->f, c->w, d, p.dt, p.theta, c->w.err);
  HIPCHK(c, hipGetLastError());
  return 0;
}

static void pcg_args(gomgpu_ctx *c, PcgArgs &a) {
  a.diag = c->w.fs_diag;
  a.off = c->w.fs_off;
  a.b = c->w.fs_rhs;
  a.col = c->m.adj_cellnum_at_cell;
  a.nfaces = c->m.tri_or_quad;
  a.x = c->w.pcg_x;
  a.r = c->w.pcg_r;
  a.d0 = c->w.pcg_d0;
  a.d1 = c->w.pcg_d1;
  a.q = c->w.pcg_q;
  a.s = c->w.pcg_s;
  a.minv = c->w.pcg_minv;
  a.partA = c->w.pcg_part;
  long pstride = c->pcg_blocks > 4096 ? c->pcg_blocks : 4096;
  a.partB = c->w.pcg_part + pstride;
  a.partC = c->w.pcg_part + 2 * pstride;
  a.partD = c->w.pcg_part + 3 * pstride;
  a.pub = c->w.pcg_pub;
  a.slots = c->w.pcg_slots;
  a.info = c->w.pcg_info;
  a.bar = c->w.pcg_bar;
  a.barrier_mode = c->pcg_barrier;
  a.n = c->d.E;
  a.maxit = c->p.max_iteration_pcg;
  a.tol = c->p.error_tolerance;
}

// The FP32 side of the mixed-precision solve, on the context's PCG buffers: the FP32 matrix
// in pcg_minv (diagonal) and the upper half of pcg_pub (4 slots), the FP32 published values in
// its lower half, the FP32 residual in pcg_r, the FP64 correction and direction in pcg_s and
// pcg_d0,
// the outer partial sums in the third partial buffer and gamma_0 in the fourth.
static void pcg_args32(gomgpu_ctx *c, const PcgArgs &a, PcgArgs32 &m) {
  const long E = c->d.E;
  const long pstride = c->pcg_blocks > 4096 ? c->pcg_blocks : 4096;
  m.diag = (const float *)c->w.pcg_minv;
  m.off = (const float *)(c->w.pcg_pub + 3 * E);
  m.b = (const float *)c->w.pcg_r;
  m.col = a.col;
  m.nfaces = a.nfaces;
  m.x = c->w.pcg_s;
  m.d0 = c->w.pcg_d0;
  m.pub = (float *)c->w.pcg_pub;
  m.partA = c->w.pcg_part;
  m.partB = c->w.pcg_part + pstride;
  m.slots = a.slots;
  m.info = a.info;
  m.bar = a.bar;
  m.n = a.n;
  m.maxit = a.maxit;
  m.tol = a.tol;
  m.gpart = c->w.pcg_part + 2 * pstride;
  m.ngpart = 0;
  m.first = 1;
  m.g0 = c->w.pcg_part + 3 * pstride;
}

static int fs_pcg(gomgpu_ctx *c) {
  const DevDims &d = c->d;
  hipStream_t st = c->st;
  PcgArgs a;
  pcg_args(c, a);
  KT_BEGIN(KT_PCG);
  if (c->p.precision == GOMGPU_MIXED) {
    PcgArgs32 m;
    pcg_args32(c, a, m);
    if (gom_pcg_cg1m_blocks(d.E) > gom_pcg_cg1m32_capacity(c->device))
      return seterr(c, GOMGPU_E_HIP, "the mixed-precision PCG grid does not fit the device");
    HIPCHK(c, gom_launch_pcg_mixed(&a, &m, c->pcg_coop, st));
  } else {
    hipError_t le = hipErrorInvalidValue;
    if (c->pcg_mode >= 1 && c->pcg_mode <= 3) {
      int rmax = c->pcg_bs >= 1024 ? 2 : c->pcg_bs >= 512 ? 4 : 8;
      int nb = c->pcg_nb > 0 ? c->pcg_nb
               : c->pcg_bs == 512 ? pcg_blocks_for(d.E, c->pcg_rows)
                                  : (int)((d.E + (long)rmax * c->pcg_bs - 1) / ((long)rmax * c->pcg_bs));
      if (nb < 1) nb = 1;
      if (nb > 4096) nb = 4096;
      if (!c->pcg_coop && nb > c->ncu) nb = c->ncu;  // plain launch: at most one block per CU
      if (c->pcg_mode == 2 || c->pcg_mode == 3) {
        int slots = c->pcg_mode == 3 && nb <= PCG_MAX_SLOT_BLOCKS;
        if (nb <= gom_pcg_cg1_capacity(c->device, d.E, nb, c->pcg_bs, slots))
          le = gom_launch_pcg_cg1(&a, nb, c->pcg_bs, slots, c->pcg_coop, st);
      } else if (nb <= gom_pcg_reg_capacity(c->device, d.E, nb, c->pcg_bs)) {
        le = gom_launch_pcg_reg(&a, nb, c->pcg_bs, c->pcg_coop, st);
      }
    }
    if (le == hipErrorInvalidValue && (c->pcg_mode == 3 || c->pcg_mode == 4)) {
      // rows beyond the register capacity: the memory-resident single-reduction solve
      int nb = gom_pcg_cg1m_blocks(d.E);
      if (nb <= gom_pcg_cg1m_capacity(c->device)) le = gom_launch_pcg_cg1m(&a, nb, c->pcg_coop, st);
    }
    if (le == hipErrorInvalidValue) {
      // memory-resident two-reduction solve; at most one block per CU (the same grid in a batch)
      int cap = c->pcg_blocks < c->ncu ? c->pcg_blocks : c->ncu;
      if (c->pcg_cap > 0 && c->pcg_cap < cap) cap = c->pcg_cap;
      le = gom_launch_pcg(&a, cap, c->pcg_coop, st);
    }
    HIPCHK(c, le);
  }
  KT_END(KT_PCG);
  c->pcg_ran = true;
  return 0;
}

static int fs_post(gomgpu_ctx *c) {
  const DevDims &d = c->d;
  const int BS = 256;
  k_copy<<<nblk(d.E, BS), BS, 0, c->st>>>(c->s.eta_cell_new, c->w.pcg_x, d.E);
  k_eta_node<<<nblk(d.P, BS), BS, 0, c->st>>>(c->m, c->s, d);
  k_eta_node_dry<<<nblk(d.P, BS), BS, 0, c->st>>>(c->m, c->s, d);
  HIPCHK(c, hipGetLastError());
  return 0;
}

#define LB(k) (c->lb[GOMGPU_LAUNCH_##k])

static int stage_impl(gomgpu_ctx *c, int id) {
  const gomgpu_params &p = c->p;
  const DevDims &d = c->d;
  const int BS = LB(STREAM);
  hipStream_t st = c->st;
  int *err = c->w.err;
  unsigned reads = 0;  // exchanged new arrays the stage reads or writes in part
  if (id == GOMGPU_STAGE_LAYERS) reads = SWAP_WN;
  if (id == GOMGPU_STAGE_NODEVEL || id == GOMGPU_STAGE_TURBULENCE) reads = SWAP_UN | SWAP_VN;
  if (id == GOMGPU_STAGE_TRANSPORT) reads = SWAP_UN | SWAP_WN;
  if (int rc = materialize(c, reads)) return rc;
  switch (id) {
    case GOMGPU_STAGE_FRICTION:
      k_friction<<<nblk(d.J, BS), BS, 0, st>>>(c->m, c->s, d, p.bf_flag, p.dry_depth, p.von_Karman);
      break;
    case GOMGPU_STAGE_ELM:
      // compute_boundary_velocity (main.f90:152) runs on the host before the hot path in
      // the reference; with river_on_device it is done here, before its first use.
      if (p.river_on_device && d.NQB > 0)
        k_river_velocity<<<nblk(d.NQB, 64), 64, 0, st>>>(c->m, c->s, c->f, d, (double *)c->f.Qu_boundary);
      if (p.advection_flag == 0) {
        long n = (long)d.L * d.J;
        k_elm_copy<<<nblk(n, BS), BS, 0, st>>>(c->s, n);
      } else {
        if (p.ELM_backtrace_flag == 1)
          k_elm_substeps<<<nblk((long)d.P * d.L, BS), BS, 0, st>>>(c->m, c->s, d, p.dt, p.delta_z_min,
                                                                   p.ELM_min_iter, p.ELM_max_iter);
        KT_BEGIN(KT_ELM);
        k_elm<<<nblk(tile_threads(d.J, d.L), LB(ELM)), LB(ELM), 0, st>>>(c->m, c->s, d, p.MSL, p.dt, p.adv_onoff_depth, err,
                                                                  c->w.elm_dep, c->w.elm_near);
        KT_END(KT_ELM);
      }
      break;
    case GOMGPU_STAGE_MOMENTUM: {
      MomParams q{p.dt, p.gravity, p.theta, p.rho_o, p.smagorinsky_parameter, p.Ah_0, p.Kh_0, p.baroclinic_flag};
      KT_BEGIN(KT_MOMENTUM);
      k_cell_uv<<<nblk((long)((d.L + GOM_CUV_LAYERS - 1) / GOM_CUV_LAYERS) * d.E, BS), BS, 0, st>>>(c->m, c->s, c->w,
                                                                                                 d);
      // 64 faces per block, all their rows; LDS for the baroclinic products of every layer
      k_mom_rhs<<<nblk(d.J, 64), LB(MOM_RHS), (size_t)2 * d.L * 64 * sizeof(double), st>>>(c->m, c->s, c->f, c->w,
                                                                                          d, q);
      k_mom_solve<<<nblk(d.J, LB(MOM_SOLVE)), LB(MOM_SOLVE), 0, st>>>(c->m, c->s, c->f, c->w, d, q, err);
      KT_END(KT_MOMENTUM);
      break;
    }
    case GOMGPU_STAGE_FREESURF: {
      int rc = fs_pre(c);
      if (!rc) rc = fs_pcg(c);
      if (!rc) rc = fs_post(c);
      if (rc) return rc;
      break;
    }
    case GOMGPU_STAGE_VELOCITY:
      k_velocity_face<<<nblk(d.J, BS), BS, 0, st>>>(c->m, c->s, c->f, d, p.theta * p.gravity * p.dt);
      k_velocity_w<<<nblk(d.E, BS), BS, 0, st>>>(c->m, c->s, d);
      c->stale &= ~(unsigned)(SWAP_UN | SWAP_VN | SWAP_WN);  // rewritten in full
      break;
    case GOMGPU_STAGE_LAYERS:
      c->layers_since_update++;
      k_layers_cell<<<nblk(d.E, BS), BS, 0, st>>>(c->m, c->s, d, p.MSL, p.dry_depth, err);
      k_layers_face<<<nblk(d.J, BS), BS, 0, st>>>(c->m, c->s, d, p.MSL, p.dry_depth, err);
      k_layers_node<<<nblk(d.P, BS), BS, 0, st>>>(c->m, c->s, d, p.MSL, p.dry_depth, err);
      k_layers_check_cell<<<nblk(d.E, BS), BS, 0, st>>>(c->m, c->s, d, p.denominator_min_for_matrix, err);
      k_layers_check_fn<<<nblk((long)d.J + d.P, BS), BS, 0, st>>>(c->m, c->s, d, p.denominator_min_for_matrix, err);
      break;
    case GOMGPU_STAGE_NODEVEL:
      k_face_levels<<<nblk(d.J, BS), BS, 0, st>>>(c->m, c->s, d);
      k_node_vel<<<nblk((long)d.P * ((d.L + GOM_NV_LEVELS) / GOM_NV_LEVELS), LB(NODE_VEL)), LB(NODE_VEL), 0, st>>>(c->m, c->s, d);
      break;
    case GOMGPU_STAGE_TRANSPORT: {
      if (p.transport_flag != 1 || d.NT == 0) break;
      if (p.transport_solver == 2) {
        // ELM transport: the backtraced, diffused right-hand sides per cell-layer, the column
        // solves, node and face means. A dry column keeps its new values, so a stale pair
        // (new standing for old) has old copied into such a column.
        const int copy_salt = (c->stale & SWAP_SALT) != 0, copy_temp = (c->stale & SWAP_TEMP) != 0;
        const int Nt = p.trans_sub_iter > 0 ? p.trans_sub_iter : 1;
        KT_BEGIN(KT_TVD_COLUMN);
        k_elm_tracer<<<nblk(tile_threads(d.E, d.L), 256), 256, 0, st>>>(c->m, c->s, c->f, c->w, d, p.MSL, p.dt, p.rho_o, Nt,
                                                                 err);
        k_elm_column<<<nblk(d.E, 64), 64, 0, st>>>(c->m, c->s, c->w, d, p.dt, copy_salt, copy_temp, err);
        KT_END(KT_TVD_COLUMN);
        k_elm_node<<<nblk(d.P, BS), BS, 0, st>>>(c->m, c->s, d);
        k_elm_face<<<nblk((long)d.J * d.L, BS), BS, 0, st>>>(c->m, c->s, d);
        c->stale &= ~(unsigned)(SWAP_SALT | SWAP_TEMP);
        break;
      }
      TvdArgs a;
      a.Nt = p.trans_sub_iter > 0 ? p.trans_sub_iter : 1;
      a.dt = p.dt;
      a.dt2 = p.dt / a.Nt;
      a.th = p.theta;
      a.rho_o = p.rho_o;
      a.hlim = p.h_flux_limiter;
      a.vlim = p.v_flux_limiter;
      a.heat = p.heat_option;
      a.sub = 0;
      long nE2 = (long)d.E * (d.L + 2);
      const bool f32 = p.precision == GOMGPU_MIXED;
      k_fill<<<nblk((long)d.NT * d.L * d.E, BS), BS, 0, st>>>(c->w.con_new, 0.0, (long)d.NT * d.L * d.E);
      tvd_launch_init(f32, nE2, BS, st, c->m, c->s, c->f, c->w, d, a);
      // the sub-cycles read con from one buffer and write the next sub-cycle's into the other
      double *cbuf[2] = {c->w.con, c->w.con2};
      for (int sub = 0; sub < a.Nt; sub++) {
        a.sub = sub;
        DevWork wv = c->w;
        wv.con = cbuf[sub & 1];
        long nb = (long)d.L * (d.NOB + d.NQB);
        if (nb > 0) tvd_launch_bnd(f32, nb, BS, st, c->m, c->s, c->f, wv, d, a);
        tvd_launch_coef_face(f32, (long)d.J * d.L, BS, st, c->m, c->s, wv, d, a);
        tvd_launch_coef_cell(f32, (long)d.E * (d.L + 1), BS, st, c->m, c->s, wv, d, a);
        if (p.h_flux_limiter > 0 || p.v_flux_limiter > 0) {
          tvd_launch_lim_cell(f32, (long)((d.L + GOM_TVD_LIM_LEVELS) / GOM_TVD_LIM_LEVELS) * d.E, LB(TVD_LIM), st, c->m, c->s, wv, d, p.h_flux_limiter,
                              sub == 0);
          tvd_launch_lim_face(f32, (long)d.J * d.L, BS, st, c->m, c->s, wv, d, p.h_flux_limiter, sub == 0);
        }
        KT_BEGIN(KT_TVD_COLUMN);
        tvd_launch_rhs(f32, (long)d.L * d.E, LB(TVD_RHS), st, c->m, c->s, wv, d, a);
        tvd_launch_solve(f32, d.E, LB(TVD_SOLVE), st, c->m, c->s, wv, d, a, cbuf[(sub + 1) & 1], sub < 2,
                         sub == a.Nt - 1, err);
        KT_END(KT_TVD_COLUMN);
      }
      tvd_launch_node(f32, (long)d.P * ((d.L + GOM_TVD_NODE_LEVELS - 1) / GOM_TVD_NODE_LEVELS), BS, st, c->m, c->s,
                      c->w, d);
      tvd_launch_face_out(f32, (long)d.J * d.L, BS, st, c->m, c->s, c->w, d);
      long nmax = (long)d.L * (d.E > d.P ? d.E : d.P);
      tvd_launch_cellnode_out(f32, nmax, BS, st, c->s, c->w, d);
      for (int t = 0; t < d.NT; t++) c->stale &= ~(unsigned)(d.tid[t] == 1 ? SWAP_SALT : SWAP_TEMP);
      break;
    }
    case GOMGPU_STAGE_EOS:
      if (p.baroclinic_flag != 1) break;
      k_eos_nodeface<<<nblk((long)d.P + d.J, 128), 128, 0, st>>>(c->m, c->s, d, p.MSL, p.ana_density == 0,
                                                                 p.i_density_flag, err);
      k_eos_cell<<<nblk(d.E, BS), BS, 0, st>>>(c->m, c->s, d, err);
      break;
    case GOMGPU_STAGE_UPDATE: {
      // update_variables.f90:16-31: new -> old. A pair whose new array every step rewrites in
      // full before reading it -- wn_cell (k_velocity_w), un_face and vn_face
      // (k_velocity_face), and each transported tracer (k_tvd_cellnode_out) -- exchanges
      // its two pointers instead of copying (the pair is then stale, see materialize). The
      // others are copied: eta_cell_new is the next solve's initial guess, and the layer
      // kernels write dz and dzhalf over the wet range only, so only that range is copied
      // (k_update_dz). Contexts in a batch copy every pair in full, since a captured batch
      // step holds their pointers and replays without these decisions.
      if (int rc = materialize(c, SWAP_ALL)) return rc;
      size_t E = d.E, J = d.J, P = d.P, L = d.L;
      const bool swap = c->batches.empty();
      unsigned swappable = SWAP_WN | SWAP_UN | SWAP_VN;
      for (int t = 0; p.transport_flag == 1 && t < d.NT; t++) swappable |= d.tid[t] == 1 ? SWAP_SALT : SWAP_TEMP;
      for (unsigned bit = 1; bit <= SWAP_TEMP; bit <<= 1) {
        double **o, **nw;
        size_t n;
        swap_pair(c, bit, &o, &nw, &n);
        if (swap && (swappable & bit)) {
          std::swap(*o, *nw);
          c->stale |= bit;
        } else {
          k_copy<<<nblk((long)n, BS), BS, 0, st>>>(*o, *nw, (int)n);
        }
      }
      k_copy<<<nblk((long)E, BS), BS, 0, st>>>(c->s.eta_cell, c->s.eta_cell_new, (int)E);
      if (!swap || c->dz_full || c->layers_since_update != 1) {  // batch members: always all
        struct {
          double *dst;
          const double *src;
          size_t n;
        } cp[] = {{c->s.dz_cell, c->s.dz_cell_new, L * E},
                  {c->s.dzhalf_cell, c->s.dzhalf_cell_new, (L + 1) * E},
                  {c->s.dz_face, c->s.dz_face_new, L * J},
                  {c->s.dzhalf_face, c->s.dzhalf_face_new, (L + 1) * J},
                  {c->s.dz_node, c->s.dz_node_new, L * P},
                  {c->s.dzhalf_node, c->s.dzhalf_node_new, (L + 1) * P}};
        for (auto &x : cp) k_copy<<<nblk((long)x.n, BS), BS, 0, st>>>(x.dst, x.src, (int)x.n);
      } else {
        k_update_dz<<<nblk((long)E, BS), BS, 0, st>>>(c->s.dz_cell, c->s.dz_cell_new, c->s.dzhalf_cell,
                                                              c->s.dzhalf_cell_new, c->s.top_layer_at_element,
                                                              c->m.bottom_layer_at_element, (int)E, (int)L);
        k_update_dz<<<nblk((long)J, BS), BS, 0, st>>>(c->s.dz_face, c->s.dz_face_new, c->s.dzhalf_face,
                                                              c->s.dzhalf_face_new, c->s.top_layer_at_face,
                                                              c->m.bottom_layer_at_face, (int)J, (int)L);
        k_update_dz<<<nblk((long)P, BS), BS, 0, st>>>(c->s.dz_node, c->s.dz_node_new, c->s.dzhalf_node,
                                                              c->s.dzhalf_node_new, c->s.top_layer_at_node,
                                                              c->m.bottom_layer_at_node, (int)P, (int)L);
      }
      c->dz_full = false;
      c->layers_since_update = 0;
      k_bars<<<nblk(d.P, BS), BS, 0, st>>>(c->m, c->s, d);
      break;
    }
    case GOMGPU_STAGE_BARS:
      // calculate_vertically_averaged_values (prepare_gom.f90:172)
      k_bars<<<nblk(d.P, BS), BS, 0, st>>>(c->m, c->s, d);
      break;
    case GOMGPU_STAGE_BOUNDARY: {
      // compute_boundary_eta, calculate_Q and the transport's boundary concentrations
      // (k_boundary.hip), from the tables of gomgpu_set_boundary
      if (!c->bnd_on) break;
      HIPCHK(c, gom_launch_boundary(&c->bnd, &c->f, &d, p.dt, c->bnd_nt, p.transport_solver == 2, st));
      break;
    }
    case GOMGPU_STAGE_WIND: {
      // read_hurricane_ser_1.f90 (node fields and face winds) and the hurricane branch of
      // calculate_wind_stress.f90:88-114; the step's phase and spin-up are device scalars
      if (p.hurricane_flag != 1) break;
      if (c->holland_on) k_holland_next<<<nblk(d.P, BS), BS, 0, st>>>(c->m, c->f, c->holland, d.P);
      k_hurricane_nodes<<<nblk(d.P, BS), BS, 0, st>>>(c->f, d.P, (double)p.hurricane_read_interval);
      k_wind_faces<<<nblk(d.J, BS), BS, 0, st>>>(c->m, c->f, d.J, p.rho_a / p.rho_o, p.wind_formular);
      break;
    }
    case GOMGPU_STAGE_TURBULENCE: {
      if (p.turbulence_flag == 0) break;
      TurbParams tp;
      tp.dt = p.dt;
      tp.g = p.gravity;
      tp.rho_o = p.rho_o;
      tp.MSL = p.MSL;
      tp.fs_coef = (double)powf(16.6f, 2.0f / 3.0f);  // 16.6**(2.0/3) in default real
      tp.q2min = p.q2min;
      tp.psimin = p.psimin;
      tp.eps_min = p.eps_min;
      tp.bgdiff = p.bgdiff;
      tp.diffmax = p.diffmax;
      tp.xlmin1 = p.xlmin1;
      tp.xlmin2 = p.xlmin2;
      tp.cmiu0 = p.cmiu0;
      tp.cpsi1 = p.cpsi1;
      tp.cpsi2 = p.cpsi2;
      tp.rpub = p.rpub;
      tp.rmub = p.rmub;
      tp.rnub = p.rnub;
      tp.schk = p.schk;
      tp.schpsi = p.schpsi;
      tp.h1_pp = p.h1_pp;
      tp.h2_pp = p.h2_pp;
      tp.vdmax_pp1 = p.vdmax_pp1;
      tp.vdmax_pp2 = p.vdmax_pp2;
      tp.vdmin_pp1 = p.vdmin_pp1;
      tp.vdmin_pp2 = p.vdmin_pp2;
      tp.tdmin_pp = p.tdmin_pp;
      tp.flag = p.turbulence_flag;
      tp.model = p.turbulence_model;
      tp.stab = p.stability_function;
      if (tp.flag == 3) {
        k_turb_column<<<nblk(d.J, LB(TURB_SOLVE)), LB(TURB_SOLVE), 0, st>>>(c->m, c->s, c->f, c->w, d, tp, err);
        k_turb_levels<<<nblk(d.J, LB(TURB_SOLVE)), LB(TURB_SOLVE), 0, st>>>(c->m, c->s, c->f, c->w, d, tp, err);
      } else {
        k_turb_ri<<<nblk((long)d.L * d.J, LB(TURB_PREP)), LB(TURB_PREP), 0, st>>>(c->m, c->s, c->w, d, tp);
      }
      k_turb_kv<<<nblk((long)((d.L + GOM_KV_LEVELS) / GOM_KV_LEVELS) * d.E, BS), BS, 0, st>>>(c->m, c->s, c->w, d);
      break;
    }
    default:
      return seterr(c, GOMGPU_E_ARG, "bad stage id");
  }
  HIPCHK(c, hipGetLastError());
  return 0;
}

extern "C" int gomgpu_stage(gomgpu_ctx *c, int id) {
  if (!c) return GOMGPU_E_ARG;
  return stage_impl(c, id);
}

extern "C" int gomgpu_step(gomgpu_ctx *c) {
  if (!c) return GOMGPU_E_ARG;
  if (c->timing) HIPCHK(c, hipEventRecord(c->ev[0], c->st));
  for (int pos = 0; pos < kNumStep; pos++) {
    int rc = stage_impl(c, kStepOrder[pos]);
    if (rc) return rc;
    if (c->timing) HIPCHK(c, hipEventRecord(c->ev[pos + 1], c->st));
  }
  if (c->timing) {
    HIPCHK(c, hipEventSynchronize(c->ev[kNumStep]));
    for (int pos = 0; pos < kNumStep; pos++) {
      float ms = 0;
      hipEventElapsedTime(&ms, c->ev[pos], c->ev[pos + 1]);
      c->stage_ms[kStepOrder[pos]] = ms;
    }
  }
  return 0;
}

extern "C" int gomgpu_run(gomgpu_ctx *c, int nsteps) {
  for (int i = 0; i < nsteps; i++) {
    int rc = gomgpu_step(c);
    if (rc) return rc;
  }
  return 0;
}

extern "C" int gomgpu_sync(gomgpu_ctx *c) {
  if (!c) return GOMGPU_E_ARG;
  HIPCHK(c, hipMemcpyAsync(c->h_err, c->w.err, 2 * sizeof(int), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(c->h_err + 2, c->w.pcg_info, 3 * sizeof(int), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  bool ran = c->pcg_ran;
  c->pcg_ran = false;
  if (ran) c->last_iters = c->h_err[2];
  if (c->h_err[0] != 0) {
    char buf[160];
    snprintf(buf, sizeof buf, "model error %d at entity %d", c->h_err[0], c->h_err[1]);
    c->err = buf;
    return GOMGPU_E_MODEL;
  }
  if (ran && c->h_err[4] != 0) {
    c->err = "PCG slot exchange timed out (blocks not co-resident)";
    return GOMGPU_E_HIP;
  }
  if (ran && c->h_err[3] == 0) {
    c->err = "PCG did not converge within max_iteration_pcg";
    return GOMGPU_W_PCG_NOT_CONVERGED;
  }
  return 0;
}

extern "C" int gomgpu_pcg_iterations(gomgpu_ctx *c) {
  if (!c) return -1;
  gomgpu_sync(c);
  return c->last_iters;
}

extern "C" int gomgpu_budgets(gomgpu_ctx *c, double *volume, double *salt) {
  if (!c || !volume || !salt) return GOMGPU_E_ARG;
  int nb = nblk(c->d.E, 256);
  k_budget<<<nb, 256, 0, c->st>>>(c->m, c->s, c->d, c->w.budget);
  HIPCHK(c, hipGetLastError());
  std::vector<double> part(2 * (size_t)nb);
  HIPCHK(c, hipMemcpyAsync(part.data(), c->w.budget, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  double v = 0.0, s = 0.0;
  for (int b = 0; b < nb; b++) {
    v += part[2 * b];
    s += part[2 * b + 1];
  }
  *volume = v;
  *salt = s;
  return 0;
}

extern "C" const char *gomgpu_last_error(gomgpu_ctx *c) {
  if (c) return c->err.c_str();
  return g_create_err.empty() ? "null context" : g_create_err.c_str();
}

extern "C" int gomgpu_timings(gomgpu_ctx *c, int enable, const char **names, double *ms, int n) {
  if (!c) return GOMGPU_E_ARG;
  bool was = c->timing;
  c->timing = enable != 0;
  int k = 0;
  for (; k < n && k < GOMGPU_NUM_STAGES; k++) {
    if (names) names[k] = kStageNames[k];
    if (ms) ms[k] = c->stage_ms[k];
  }
  for (int i = 0; i < GOMGPU_NUM_KTIMERS && k < n; i++, k++) {
    if (names) names[k] = kKernelNames[i];
    if (ms) ms[k] = c->kernel_calls[i] ? c->kernel_ms[i] / c->kernel_calls[i] : 0.0;
  }
  if (c->timing && !was)
    for (int i = 0; i < GOMGPU_NUM_KTIMERS; i++) {
      c->kernel_ms[i] = 0.0;
      c->kernel_calls[i] = 0;
    }
  return k;
}

// ------------------------------------------------------------------ batches
// A batch advances several independent scenarios (contexts on one device) by one step in
// one submission. Per member the stages run on the member's own stream, so small members
// overlap on the device; the PCG solves of all members run in fused launches between a
// join and a fork of those streams (one k_pcg_cg1_batch launch per group of members whose
// blocks fit the device together; a member too large for the register-resident solve gets
// the memory-resident k_pcg on the batch stream). With graphs on, the whole submission is
// captured once per membership and replayed every step.
extern "C" hipError_t gom_launch_pcg_cg1_batch(const PcgArgs *d_members, const int *d_base, const int *d_nbm,
                                               int nmem, int total_blocks, hipStream_t st);
extern "C" int gom_pcg_cg1_batch_capacity(int device);

struct BatchGroup {
  int nmem = 0, total = 0;
  PcgArgs *d_args = nullptr;
  int *d_base = nullptr, *d_nbm = nullptr;
};

struct gomgpu_batch {
  int device = 0;
  hipStream_t st = nullptr;
  std::vector<gomgpu_ctx *> members;
  std::vector<hipEvent_t> ev_join;  // one per member, reused
  hipEvent_t ev_fork = nullptr, ev_done = nullptr;
  std::vector<BatchGroup> groups;
  std::vector<gomgpu_ctx *> solo;
  bool dirty = true, use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int capacity = 0;
  std::string err;
};

static int bseterr(gomgpu_batch *b, int code, const std::string &msg) {
  b->err = msg;
  return code;
}
#define BHIP(b, call)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) return bseterr(b, GOMGPU_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

static void batch_free_plan(gomgpu_batch *b) {
  for (auto &g : b->groups) {
    if (g.d_args) hipFree(g.d_args);
    if (g.d_base) hipFree(g.d_base);
    if (g.d_nbm) hipFree(g.d_nbm);
  }
  b->groups.clear();
  b->solo.clear();
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

static int batch_plan(gomgpu_batch *b) {
  batch_free_plan(b);
  std::vector<std::vector<gomgpu_ctx *>> grp;
  int total = 0;
  for (gomgpu_ctx *c : b->members) {
    int nb = pcg_blocks_for(c->d.E, c->pcg_rows);
    if (nb > b->capacity || nb > PCG_MAX_SLOT_BLOCKS || c->p.precision == GOMGPU_MIXED) {
      b->solo.push_back(c);
      continue;
    }
    int padded = (nb + 7) / 8 * 8;
    if (grp.empty() || total + padded > b->capacity) {
      grp.emplace_back();
      total = 0;
    }
    grp.back().push_back(c);
    total += padded;
  }
  for (auto &gm : grp) {
    BatchGroup g;
    std::vector<PcgArgs> args;
    std::vector<int> base, nbm;
    int off = 0;
    for (gomgpu_ctx *c : gm) {
      PcgArgs a;
      pcg_args(c, a);
      args.push_back(a);
      int nb = pcg_blocks_for(c->d.E, c->pcg_rows);
      base.push_back(off);
      nbm.push_back(nb);
      off += (nb + 7) / 8 * 8;
    }
    g.nmem = (int)gm.size();
    g.total = off;
    BHIP(b, hipMalloc(&g.d_args, sizeof(PcgArgs) * args.size()));
    BHIP(b, hipMalloc(&g.d_base, sizeof(int) * base.size()));
    BHIP(b, hipMalloc(&g.d_nbm, sizeof(int) * nbm.size()));
    BHIP(b, hipMemcpyAsync(g.d_args, args.data(), sizeof(PcgArgs) * args.size(), hipMemcpyHostToDevice, b->st));
    BHIP(b, hipMemcpyAsync(g.d_base, base.data(), sizeof(int) * base.size(), hipMemcpyHostToDevice, b->st));
    BHIP(b, hipMemcpyAsync(g.d_nbm, nbm.data(), sizeof(int) * nbm.size(), hipMemcpyHostToDevice, b->st));
    BHIP(b, hipStreamSynchronize(b->st));
    b->groups.push_back(g);
  }
  return 0;
}

static const int kPreStages[] = {GOMGPU_STAGE_BOUNDARY, GOMGPU_STAGE_WIND, GOMGPU_STAGE_FRICTION, GOMGPU_STAGE_ELM,
                                 GOMGPU_STAGE_MOMENTUM};
static const int kPostStages[] = {GOMGPU_STAGE_VELOCITY,   GOMGPU_STAGE_LAYERS,    GOMGPU_STAGE_NODEVEL,
                                  GOMGPU_STAGE_TURBULENCE, GOMGPU_STAGE_TRANSPORT, GOMGPU_STAGE_EOS,
                                  GOMGPU_STAGE_UPDATE};

// The submission of one batch step on b->st (captured or run eagerly).
static int batch_enqueue(gomgpu_batch *b) {
  BHIP(b, hipEventRecord(b->ev_fork, b->st));
  for (size_t i = 0; i < b->members.size(); i++) {
    gomgpu_ctx *c = b->members[i];
    BHIP(b, hipStreamWaitEvent(c->st, b->ev_fork, 0));
    for (int id : kPreStages)
      if (int rc = stage_impl(c, id)) return bseterr(b, rc, c->err);
    if (int rc = fs_pre(c)) return bseterr(b, rc, c->err);
    BHIP(b, hipEventRecord(b->ev_join[i], c->st));
    BHIP(b, hipStreamWaitEvent(b->st, b->ev_join[i], 0));
  }
  for (auto &g : b->groups) BHIP(b, gom_launch_pcg_cg1_batch(g.d_args, g.d_base, g.d_nbm, g.nmem, g.total, b->st));
  for (gomgpu_ctx *c : b->solo) {  // the solve each of them runs alone (fs_pcg)
    PcgArgs a;
    pcg_args(c, a);
    if (c->p.precision == GOMGPU_MIXED) {
      PcgArgs32 m;
      pcg_args32(c, a, m);
      BHIP(b, gom_launch_pcg_mixed(&a, &m, 0, b->st));
      continue;
    }
    int nb = gom_pcg_cg1m_blocks(c->d.E);
    if (nb <= gom_pcg_cg1m_capacity(c->device)) {
      BHIP(b, gom_launch_pcg_cg1m(&a, nb, 0, b->st));
    } else {
      int cap = c->pcg_blocks < c->ncu ? c->pcg_blocks : c->ncu;
      BHIP(b, gom_launch_pcg(&a, cap, 0, b->st));
    }
  }
  BHIP(b, hipEventRecord(b->ev_fork, b->st));
  for (size_t i = 0; i < b->members.size(); i++) {
    gomgpu_ctx *c = b->members[i];
    BHIP(b, hipStreamWaitEvent(c->st, b->ev_fork, 0));
    if (int rc = fs_post(c)) return bseterr(b, rc, c->err);
    for (int id : kPostStages)
      if (int rc = stage_impl(c, id)) return bseterr(b, rc, c->err);
    BHIP(b, hipEventRecord(b->ev_join[i], c->st));
    BHIP(b, hipStreamWaitEvent(b->st, b->ev_join[i], 0));
  }
  return 0;
}

static int batch_capture(gomgpu_batch *b) {
  BHIP(b, hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
  int rc = batch_enqueue(b);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->st, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return bseterr(b, GOMGPU_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  b->graph = g;
  BHIP(b, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  return 0;
}

extern "C" int gomgpu_batch_create(gomgpu_batch **out) {
  if (!out) return GOMGPU_E_ARG;
  *out = nullptr;
  gomgpu_batch *b = new gomgpu_batch();
  if (hipGetDevice(&b->device) != hipSuccess || hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_done, hipEventDisableTiming) != hipSuccess) {
    if (b->st) hipStreamDestroy(b->st);
    delete b;
    return GOMGPU_E_HIP;
  }
  b->capacity = gom_pcg_cg1_batch_capacity(b->device);
  if (const char *e = getenv("GOMGPU_BATCH_GRAPH")) b->use_graph = atoi(e) != 0;
  *out = b;
  return 0;
}

extern "C" int gomgpu_batch_add(gomgpu_batch *b, gomgpu_ctx *c) {
  if (!b || !c) return GOMGPU_E_ARG;
  if (c->device != b->device) return bseterr(b, GOMGPU_E_ARG, "member on another device");
  for (gomgpu_ctx *x : b->members)
    if (x == c) return bseterr(b, GOMGPU_E_ARG, "context already in the batch");
  hipEvent_t ev;
  BHIP(b, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  c->timing = false;
  c->batches.push_back(b);
  b->members.push_back(c);
  b->ev_join.push_back(ev);
  b->dirty = true;
  return 0;
}

// Drops c from b (after b's queued work): the next batch step plans and captures anew.
static void batch_forget(gomgpu_batch *b, gomgpu_ctx *c) {
  if (b->st) hipStreamSynchronize(b->st);
  for (size_t i = 0; i < b->members.size(); i++)
    if (b->members[i] == c) {
      hipEventDestroy(b->ev_join[i]);
      b->members.erase(b->members.begin() + i);
      b->ev_join.erase(b->ev_join.begin() + i);
      b->dirty = true;
      break;
    }
  for (size_t i = 0; i < c->batches.size(); i++)
    if (c->batches[i] == b) {
      c->batches.erase(c->batches.begin() + i);
      break;
    }
}

extern "C" int gomgpu_batch_remove(gomgpu_batch *b, gomgpu_ctx *c) {
  if (!b || !c) return GOMGPU_E_ARG;
  for (size_t i = 0; i < b->members.size(); i++)
    if (b->members[i] == c) {
      batch_forget(b, c);
      return 0;
    }
  return bseterr(b, GOMGPU_E_ARG, "context not in the batch");
}

extern "C" int gomgpu_batch_size(gomgpu_batch *b) { return b ? (int)b->members.size() : -1; }

extern "C" int gomgpu_batch_layout(gomgpu_batch *b, int *groups, int *solo) {
  if (!b) return GOMGPU_E_ARG;
  if (groups) *groups = (int)b->groups.size();
  if (solo) *solo = (int)b->solo.size();
  return 0;
}

extern "C" int gomgpu_batch_step(gomgpu_batch *b) {
  if (!b) return GOMGPU_E_ARG;
  if (b->members.empty()) return 0;
  if (b->dirty) {
    BHIP(b, hipStreamSynchronize(b->st));
    if (int rc = batch_plan(b)) return rc;
    if (b->use_graph)
      if (int rc = batch_capture(b)) return rc;
    b->dirty = false;
  }
  // the members' own queued work (forcing uploads) precedes the step
  for (size_t i = 0; i < b->members.size(); i++) {
    BHIP(b, hipEventRecord(b->ev_join[i], b->members[i]->st));
    BHIP(b, hipStreamWaitEvent(b->st, b->ev_join[i], 0));
  }
  if (b->use_graph) {
    BHIP(b, hipGraphLaunch(b->exec, b->st));
  } else if (int rc = batch_enqueue(b)) {
    return rc;
  }
  // and the members' later work (downloads, next forcing) follows it
  BHIP(b, hipEventRecord(b->ev_done, b->st));
  for (gomgpu_ctx *c : b->members) {
    BHIP(b, hipStreamWaitEvent(c->st, b->ev_done, 0));
    c->pcg_ran = true;
  }
  return 0;
}

extern "C" int gomgpu_batch_sync(gomgpu_batch *b) {
  if (!b) return GOMGPU_E_ARG;
  BHIP(b, hipStreamSynchronize(b->st));
  int first = 0;
  for (size_t i = 0; i < b->members.size(); i++) {
    int rc = gomgpu_sync(b->members[i]);
    if (rc && !first) {
      first = rc;
      b->err = "member " + std::to_string(i) + ": " + b->members[i]->err;
    }
  }
  return first;
}

extern "C" const char *gomgpu_batch_last_error(gomgpu_batch *b) { return b ? b->err.c_str() : "null batch"; }

extern "C" void gomgpu_batch_destroy(gomgpu_batch *b) {
  if (!b) return;
  if (b->st) hipStreamSynchronize(b->st);
  batch_free_plan(b);
  for (gomgpu_ctx *c : b->members)
    for (size_t i = 0; i < c->batches.size(); i++)
      if (c->batches[i] == b) {
        c->batches.erase(c->batches.begin() + i);
        break;
      }
  for (hipEvent_t e : b->ev_join) hipEventDestroy(e);
  if (b->ev_fork) hipEventDestroy(b->ev_fork);
  if (b->ev_done) hipEventDestroy(b->ev_done);
  if (b->st) hipStreamDestroy(b->st);
  delete b;
}

// ------------------------------------------------------------------ footprint
// Device bytes gomgpu_create allocates for these parameters (the same list as create, one
// line per allocation; checked against gomgpu_device_bytes by the tests).
extern "C" long long gomgpu_footprint(const gomgpu_params *params) {
  if (!params) return -1;
  std::string why;
  if (validate(params, why)) return -1;
  const gomgpu_params &p = *params;
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) return -1;
  size_t tot = 0;
  auto add = [&](size_t count, size_t elem) { tot += (count == 0 ? 1 : count) * elem; };
#define FP_R(name, ext)                     \
  {                                         \
    Shape sh = shape_of(#ext, p);           \
    add((size_t)(sh.ld * sh.n), sizeof(double)); \
  }
#define FP_I(name, ext)                     \
  {                                         \
    Shape sh = shape_of(#ext, p);           \
    add((size_t)(sh.ld * sh.n), sizeof(int32_t)); \
  }
  GOMGPU_MESH_REAL(FP_R)
  GOMGPU_MESH_INT(FP_I)
  add(p.maxface, sizeof(double));  // Cdb_static
  {  // node stencil
    size_t P = p.maxnod, NB = p.max_no_neighbor_node;
    add(2 * NB * P, sizeof(int32_t)); add(2 * NB * P, sizeof(int32_t));
    add(NB * P, sizeof(int32_t)); add(NB * P, sizeof(int32_t));
    add(2 * NB * P, sizeof(double)); add(2 * NB * P, sizeof(double));
    add(NB * P, sizeof(double)); add(P, sizeof(double));
  }
  add(4 * (size_t)p.maxele, sizeof(double)); add(4 * (size_t)p.maxele, sizeof(double));  // ELM cell records
  add(4 * (size_t)p.maxele, sizeof(int32_t)); add(4 * (size_t)p.maxele, sizeof(int32_t));
  add(4 * (size_t)p.maxele, sizeof(int32_t));  // slot_face
  add(2 * (size_t)p.maxface, sizeof(int32_t));  // face_adj
  GOMGPU_STATE_REAL(FP_R)
  GOMGPU_STATE_INT(FP_I)
  GOMGPU_FORCING_REAL(FP_R)
#undef FP_R
#undef FP_I
  add(GOM_NSCAL, sizeof(double));  // per-step scalars
  {  // device image of the forcing staging buffer
    size_t st = GOM_NSCAL;
#define FP_S(name, ext) st += (size_t)(shape_of(#ext, p).ld * shape_of(#ext, p).n);
    GOMGPU_FORCING_REAL(FP_S)
#undef FP_S
    add(st, sizeof(double));
  }
  long E = p.maxele, J = p.maxface, P = p.maxnod, L = p.maxlayer;
  long NT = (p.transport_flag == 1 ? p.maxtran2 : 0);
  NT = NT > 0 ? NT : 1;
  long OB = p.num_ob_cell > 0 ? p.num_ob_cell : 1, QB = p.num_Qb_cell > 0 ? p.num_Qb_cell : 1;
  long pb = gom_pcg_max_blocks(device);
  const size_t D = sizeof(double);
  add(E, D); add(4 * E, D); add(E, D);                      // fs_diag, fs_off, fs_rhs
  for (int i = 0; i < 7; i++) add(E, D);                    // pcg_x .. pcg_minv
  add(4 * (pb > 4096 ? pb : 4096), D);                      // pcg_part
  add(6 * E, D); add(6 * PCG_MAX_SLOT_BLOCKS, sizeof(unsigned long long));
  add(3, sizeof(int)); add(512, sizeof(unsigned));
  add(J, D); add(J, D); add(J, D);                          // face_sdz, face_sun, face_sg
  add(NT * (L + 2) * E, D); add(NT * L * E, D);             // con, con_new
  add(2 * L * J, D); add(2 * L * J, D); add(L * J, D);      // fqd, fps, phij
  add((L + 1) * E, D); add((L + 1) * E, D); add((L + 1) * E, D);  // Qi, di, phii
  add(NT * 2 * E, D); add(NT * (L + 2) * E, D);             // htop, con2
  add(2 * (L + 1) * E, D);                                  // PSIi
  add(2 * NT * L * E, D);                                   // s12
  add(NT * L * OB, D); add(NT * L * QB, D);                 // obk, qbk
  add(NT * L * P, D); add(NT * L * J, D);                   // node_new, face_new
  add(2 * ((E + 255) / 256) + 2, D); add(2, sizeof(int));   // budget, err
  add(2, sizeof(unsigned long long));                       // elm_near
  add((L + 1) * J, D); add(L * (E > J ? E : J), D); add(NT * L * E, D);  // turb_kvf, tri_gam, tvd_x
  add(L * E, D); add(L * E, D);                                           // cell_u, cell_v
  if (p.turbulence_flag == 3)
    for (int i = 0; i < 4; i++) add(L * J, D);  // tk_shear, tk_rz, tk_q2tmp, tk_x
  return (long long)tot;
}

extern "C" long long gomgpu_device_bytes(gomgpu_ctx *c) { return c ? (long long)c->bytes : -1; }

// ------------------------------------------------------------------ launch shapes
static int launch_max(int which) {
  return which == GOMGPU_LAUNCH_TURB_SOLVE || which == GOMGPU_LAUNCH_MOM_RHS || which == GOMGPU_LAUNCH_ELM ? 256
                                                                                                     : 1024;
}

extern "C" int gomgpu_set_launch(gomgpu_ctx *c, int which, int block) {
  if (!c || which < 0 || which >= GOMGPU_NUM_LAUNCH) return GOMGPU_E_ARG;
  if (block < 64 || block > launch_max(which) || (block & (block - 1)) != 0)
    return seterr(c, GOMGPU_E_ARG, "block size must be a power of two in [64, max]");
  c->lb[which] = block;
  return 0;
}

extern "C" int gomgpu_elm_record(gomgpu_ctx *c, int record) {
  if (!c) return GOMGPU_E_ARG;
  if (record && !c->dep_buf) {
    int rc = dalloc(c, &c->dep_buf, 2 * (size_t)c->d.L * c->d.J);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->st));
  }
  int32_t *want = record ? c->dep_buf : nullptr;
  if (want != c->w.elm_dep) {
    c->w.elm_dep = want;
    for (gomgpu_batch *b : c->batches) b->dirty = true;  // a captured step holds the pointer
  }
  return 0;
}

extern "C" int gomgpu_elm_departures(gomgpu_ctx *c, int32_t *nnel, int32_t *jlev) {
  if (!c) return GOMGPU_E_ARG;
  if (!c->dep_buf) return seterr(c, GOMGPU_E_ARG, "ELM departures are not recorded (gomgpu_elm_record)");
  HIPCHK(c, hipStreamSynchronize(c->st));  // leaves the latched status to gomgpu_sync
  const size_t n = (size_t)c->d.L * c->d.J;
  if (nnel) HIPCHK(c, hipMemcpy(nnel, c->dep_buf, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (jlev) HIPCHK(c, hipMemcpy(jlev, c->dep_buf + n, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int gomgpu_elm_near_edge(gomgpu_ctx *c, long long *counts, int reset) {
  if (!c || !counts) return GOMGPU_E_ARG;
  HIPCHK(c, hipStreamSynchronize(c->st));
  unsigned long long h[2];
  HIPCHK(c, hipMemcpy(h, c->w.elm_near, sizeof h, hipMemcpyDeviceToHost));
  counts[0] = (long long)h[0];
  counts[1] = (long long)h[1];
  if (reset) {
    const unsigned long long zero[2] = {0, 0};
    HIPCHK(c, hipMemcpy(c->w.elm_near, zero, sizeof zero, hipMemcpyHostToDevice));
  }
  return 0;
}

extern "C" int gomgpu_set_pcg_rows(gomgpu_ctx *c, int rows) {
  if (!c) return GOMGPU_E_ARG;
  if (rows != 0 && rows != 1 && rows != 2 && rows != 4)
    return seterr(c, GOMGPU_E_ARG, "PCG rows per thread must be 0 (automatic), 1, 2 or 4");
  c->pcg_rows = rows;
  for (gomgpu_batch *b : c->batches) b->dirty = true;  // its batches re-plan
  return 0;
}

extern "C" int gomgpu_get_launch(gomgpu_ctx *c, int which) {
  if (!c || which < 0 || which >= GOMGPU_NUM_LAUNCH) return -1;
  return c->lb[which];
}
