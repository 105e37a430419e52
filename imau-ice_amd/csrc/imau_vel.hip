// imau_vel.hip -- host side of libimauvel.so: the C ABI of include/imau_vel.h,
// device-resident problems (one grid or an uneven-size ensemble), and the
// Picard driver of solve_DIVA / solve_SSA / solve_SIA
// (src/ice_velocity_module.f90:36-374).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/imau_net.h"
#include "../../include/imau_vel.h"
#include "dev_common.h"
#include "kernels.h"
#include "mg.h"
#include "net.h"
#include "producers.h"
#include "thickness.h"

namespace {

struct Err {
  int code;
  std::string msg;
};

#define HIPCHK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) throw Err{IMAU_ERR_DEVICE, std::string(#x ": ") + hipGetErrorString(e_)}; \
  } while (0)

[[noreturn]] void fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw Err{code, buf};
}

// vertical_integrate, src/utilities_module.f90:73-96
double vint(int nz, const double *z, const double *f) {
  double s = 0.0;
  for (int k = 1; k < nz; ++k) s += 0.5 * (f[k] + f[k - 1]) * (z[k] - z[k - 1]);
  return s;
}

template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    release();
    n = count;
    if (count) {
      HIPCHK(hipMalloc((void **)&p, count * sizeof(T)));
      HIPCHK(hipMemset(p, 0, count * sizeof(T)));
      // the zeroing runs on the null stream, which does not order with the
      // context's non-blocking stream: wait for it, so that kernels enqueued
      // next (buffers allocated on first use inside a solve) are not
      // overwritten by it
      HIPCHK(hipDeviceSynchronize());
    }
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DBuf() { release(); }
};

// one coarse level of the aggregation multigrid (explicit staggered stencils)
struct MGLevel {
  std::vector<CMember> mem;
  std::vector<Tile> tiles;
  long long na = 0;
  DBuf<CMember> mem_d;
  DBuf<Tile> tiles_d;
  DBuf<double> data;  // [cu 9na][cv 9na][xu][xv][bu][bv][x2u][x2v] (na each), FP64 or FP32 elements
};

}  // namespace

// ---------------------------------------------------------------------------
struct imau_problem {
  int ngrids = 0, nz = 0;
  std::vector<Member> mem;
  std::vector<Tile> tiles;
  std::vector<int> tstart;
  std::vector<int> sheet_any;  // member has grounded ice (else trivial solve)
  long long na = 0;            // total 2D slab size

  DBuf<Member> mem_d;
  DBuf<Tile> tiles_d;
  DBuf<int> tstart_d;
  DBuf<double> xg, yg;
  DBuf<double> ismip_e;          // ISMIP-HOM E beta per global column (offx + i)
  std::vector<char> has_slides;  // member g was created with ISMIP_HOM_E_slides
  DBuf<double> f2;    // all double 2D fields (arena)
  DBuf<int32_t> i2;   // int 2D fields
  DBuf<uint8_t> rt;   // row types
  DBuf<double> f3;    // 3D fields (arena)
  DBuf<double> stage;  // host-layout staging for 3D transfers
  DBuf<double> partial;
  DBuf<double> msum;
  DBuf<unsigned long long> resmax;
  DBuf<int> act_pic, act_lin;
  DBuf<double> sc;    // per-member Krylov scalars
  DBuf<double> relax_d;  // per-member Picard relaxation (adaptive mode, restarted members)
  DBuf<int> restart_d;   // per-member flag: reset to the created state (stall restart)
  DBuf<int> stop_d;      // per-member flag: stopped at this iteration's residual check
  DBuf<int> lin_d;       // per-member flag of the linear solve's best-iterate copies
  // DIVA_err_cx / cy (when requested): two slots [u (na)][v (na)]; Picard
  // iteration it writes slot it & 1, slot 0 starts at 1e5, and err_slot[g]
  // is the slot of member g's last estimate kept as output
  DBuf<double> err;
  std::vector<int> err_slot;
  DBuf<double> gcr_dirs; // GCR directions Q_i, Z_i (allocated on first use)
  DBuf<double> and_slots;  // Anderson slots U_s, F_s (allocated on first use)
  DBuf<double> and_w;      // Anderson weights per member
  DBuf<float> visc32;      // mixed precision: two FP32 viscosity buffers and the FP32 rate-factor term
  std::vector<std::unique_ptr<MGLevel>> mg;  // coarse levels 1.. of the multigrid preconditioner
  DBuf<int> mg_depth;                        // per member: its coarsest level
  DBuf<float> mg_f32;                        // FP32 fine-level coefficients, rhs and iterates
  DBuf<uint8_t> mg_plain;                    // level-1 plain-aggregate flags
  DBuf<int2> mg_list;                        // level-1 cells that are not plain
  DBuf<int> mg_nlist;
  DBuf<double> mg_z64;                       // FP64 scratch iterate of the fused smoother
  DBuf<Tile> mg_ftiles;                      // tiles of the fused smoother
  DBuf<Tile> vtiles;                         // IMAU_VTI-row tiles of the 3D viscosity kernel
  int nvtiles = 0;
  DBuf<Tile> mtiles;                         // tiles of its row-marching form (mti rows each)
  int nmtiles = 0, mti = IMAU_MTI;
  imau::MGPlan plan;                         // device view of the levels

  Fields F{};
  OutF O{};
  double *Acx3 = nullptr, *Acy3 = nullptr;
  double *visc_cur = nullptr, *visc_oth = nullptr;
  // uploaded state, for reset_state
  double *u0 = nullptr, *v0 = nullptr, *tbx0 = nullptr, *tby0 = nullptr, *visc0 = nullptr;
  double *u3a0 = nullptr, *v3a0 = nullptr;
  bool has_ti = false;
  int dyn_uploaded = 0;
  // per-member results
  std::vector<int> picard_its;
  std::vector<int> restarts;
  std::vector<long long> sweeps;
  std::vector<double> resid_uv, nl_resid;
  double seconds = 0.0;
  // per-call linear-solver settings (ice%DIVA_SOR_*, ice%DIVA_PETSc_*) of member 0
  int sor_nit_ = 10000;
  double sor_tol_ = 2.5, sor_omega_ = 1.3, petsc_rtol_ = 0.01, petsc_abstol_ = 2.5;
};

// a 2D download collected by imau_vel_problem_download (flush_down2)
struct Down2Req {
  int g;
  int kind;
  void *dst;
  const char *src_base;
  size_t esz;
};

struct imau_ctx {
  int device = 0;
  hipStream_t s = nullptr;
  std::vector<Down2Req> *batch = nullptr;  // non-null: down2 collects instead of copying
  void *pin[2] = {nullptr, nullptr};       // pinned staging of batched downloads, one field each
  size_t pin_bytes = 0;
  hipEvent_t pin_ev[2] = {nullptr, nullptr};
  double *scratch = nullptr;   // this device's copy of `scratch` (imau_vel_problem_residual_noalloc)
  std::string err;
  imau_problem *cached = nullptr;
  // optional HIP-event timing of the sweep kernel on this context's stream
  int ktime = 0;
  std::vector<hipEvent_t> ev;      // pairs (start, stop)
  size_t ev_used = 0;
  long long sweep_launches = 0;
  double sweep_ms = 0.0;
  long long sweep_cells = 0;
};

static std::string g_init_err;

// ---------------------------------------------------------------------------
// Scratch of imau_vel_problem_residual_noalloc, the entry point for callers
// that cannot allocate device memory: per-tile partial sums, per-member sums
// and activity flags.  Every GPU the library runs on holds its own copy until
// the process ends.  Host code reaches it only through the address
// hipGetSymbolAddress returns with that device current (scratch_address),
// never through the array's name.
constexpr size_t kScratchDoubles = IMAU_NOALLOC_SCRATCH_BYTES / sizeof(double);
static __device__ __align__(256) double scratch[kScratchDoubles];

static std::mutex g_scratch_mu;
static std::map<int, double *> g_scratch_addr;

// the address of `scratch` on `device`, which must be the current device;
// looked up once per device
static double *scratch_address(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_addr.find(device);
  if (it != g_scratch_addr.end()) return it->second;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    fail(IMAU_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
  g_scratch_addr[device] = static_cast<double *>(p);
  return static_cast<double *>(p);
}

// ---------------------------------------------------------------------------
// problem construction

// tiles of the row-marching viscosity kernel: `rows` rows x 64 columns (the
// tile's row count travels in Tile::pad); rows = 0 leaves none, so the tiled
// kernel runs.  The buffer is sized for the smallest candidate height (4), so
// retiling never allocates
static void set_mtiles(imau_problem *P, int rows) {
  std::vector<Tile> mt;
  if (rows > 0)
    for (int g = 0; g < P->ngrids; ++g)
      for (int i0 = 0; i0 < P->mem[g].nx; i0 += rows)
        for (int j0 = 0; j0 < P->mem[g].ny; j0 += IMAU_TJ) mt.push_back(Tile{g, i0, j0, rows});
  if (!P->mtiles.p) {
    size_t most = 0;
    for (int g = 0; g < P->ngrids; ++g)
      most += (size_t)((P->mem[g].nx + 3) / 4) * ((P->mem[g].ny + IMAU_TJ - 1) / IMAU_TJ);
    P->mtiles.alloc(std::max(most, mt.size()));
  }
  if (!mt.empty()) HIPCHK(hipMemcpy(P->mtiles.p, mt.data(), mt.size() * sizeof(Tile), hipMemcpyHostToDevice));
  P->nmtiles = (int)mt.size();
  P->mti = rows;
}

// Coarse levels: every level halves each component's extent (2 x 2
// aggregates, rounded up).  Member g is coarsened until it fits
// IMAU_CMAX x IMAU_CMAX (at least once); that level, depth[g], is solved by
// one workgroup in LDS, so each member's cycle is the same alone or in a batch.
static void build_mg(imau_problem *P) {
  std::vector<CMember> cur(P->ngrids);
  std::vector<int> depth(P->ngrids, 0);
  for (int g = 0; g < P->ngrids; ++g) {
    const Member &m = P->mem[g];
    cur[g] = CMember{m.nx - 1, m.ny, m.nx, m.ny - 1, m.nx, m.ny, 0};
  }
  auto fits = [](const CMember &m) { return m.nx <= IMAU_CMAX && m.ny <= IMAU_CMAX; };
  for (int lev = 1;; ++lev) {
    bool need = false;
    for (int g = 0; g < P->ngrids; ++g) need = need || depth[g] == 0;
    if (!need) break;
    auto lv = std::make_unique<MGLevel>();
    long long off = 0;
    for (int g = 0; g < P->ngrids; ++g) {
      CMember c = cur[g];
      CMember n{(c.nxu + 1) / 2, (c.nyu + 1) / 2, (c.nxv + 1) / 2, (c.nyv + 1) / 2, 0, 0, off};
      n.nx = std::max(n.nxu, n.nxv);
      n.ny = std::max(n.nyu, n.nyv);
      lv->mem.push_back(n);
      if (depth[g] == 0)
        for (int i0 = 0; i0 < n.nx; i0 += IMAU_TI)
          for (int j0 = 0; j0 < n.ny; j0 += IMAU_TJ) lv->tiles.push_back(Tile{g, i0, j0, 0});
      off += ((long long)n.nx * n.ny + 63) / 64 * 64;
      cur[g] = n;
      if (depth[g] == 0 && fits(n)) depth[g] = lev;
    }
    lv->na = off;
    lv->mem_d.alloc(lv->mem.size());
    lv->tiles_d.alloc(lv->tiles.size());
    lv->data.alloc((size_t)24 * off);
    HIPCHK(hipMemcpy(lv->mem_d.p, lv->mem.data(), lv->mem.size() * sizeof(CMember), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(lv->tiles_d.p, lv->tiles.data(), lv->tiles.size() * sizeof(Tile), hipMemcpyHostToDevice));
    P->plan.levels.push_back(imau::MGLevelDev{lv->mem_d.p, lv->tiles_d.p, (int)lv->tiles.size(), lev, off,
                                              lv->data.p, false});
    P->mg.push_back(std::move(lv));
  }
  P->mg_depth.alloc(P->ngrids);
  HIPCHK(hipMemcpy(P->mg_depth.p, depth.data(), P->ngrids * sizeof(int), hipMemcpyHostToDevice));
  for (int d : depth) P->plan.levels[d - 1].ends = true;
  P->plan.depth = P->mg_depth.p;
  P->plan.ngrids = P->ngrids;
  P->mg_plain.alloc((size_t)P->plan.levels[0].na);
  P->plan.plain = P->mg_plain.p;
  P->mg_list.alloc((size_t)P->plan.levels[0].na);
  P->mg_nlist.alloc(1);
  P->plan.list = P->mg_list.p;
  P->plan.nlist_d = P->mg_nlist.p;
  P->plan.nlist = 0;
  P->plan.coarsest_sweeps = 10;
  const long long na = P->na;
  P->mg_f32.alloc((size_t)12 * na);
  float *f = P->mg_f32.p;
  P->plan.f32 = imau::MGFine32{f, f + na, f + 2 * na, f + 3 * na, f + 4 * na, f + 5 * na,
                               f + 6 * na, f + 7 * na, f + 8 * na, f + 9 * na, f + 10 * na, f + 11 * na};
  P->mg_z64.alloc((size_t)2 * na);
  P->plan.zu64 = P->mg_z64.p;
  P->plan.zv64 = P->mg_z64.p + na;
  std::vector<Tile> ft;
  for (int g = 0; g < P->ngrids; ++g)
    for (int i0 = 0; i0 < P->mem[g].nx; i0 += IMAU_FTI)
      for (int j0 = 0; j0 < P->mem[g].ny; j0 += IMAU_FTJ) ft.push_back(Tile{g, i0, j0, 0});
  std::vector<Tile> vt;
  for (int g = 0; g < P->ngrids; ++g)
    for (int i0 = 0; i0 < P->mem[g].nx; i0 += IMAU_VTI)
      for (int j0 = 0; j0 < P->mem[g].ny; j0 += IMAU_TJ) vt.push_back(Tile{g, i0, j0, 0});
  P->vtiles.alloc(vt.size());
  HIPCHK(hipMemcpy(P->vtiles.p, vt.data(), vt.size() * sizeof(Tile), hipMemcpyHostToDevice));
  P->nvtiles = (int)vt.size();
  set_mtiles(P, IMAU_MTI);
  P->mg_ftiles.alloc(ft.size());
  HIPCHK(hipMemcpy(P->mg_ftiles.p, ft.data(), ft.size() * sizeof(Tile), hipMemcpyHostToDevice));
  P->plan.ftiles = P->mg_ftiles.p;
  P->plan.nftiles = (int)ft.size();
}

// ---------------------------------------------------------------------------
namespace {

// named 2D double fields in the arena
enum F2 {
  F_Hi, F_Hb, F_SL, F_Hs, F_TAF, F_dHs_dx_a, F_dHs_dy_a, F_dHi_dx_a, F_dHi_dy_a, F_dHb_dt, F_A_vav,
  F_Hi_cx, F_dHs_dx_cx, F_dHs_dy_cx, F_Hi_cy, F_dHs_dx_cy, F_dHs_dy_cy, F_Hi_b,
  F_beta_sq, F_alpha_sq, F_phi_fric, F_tauc,
  F_u, F_v, F_taub_cx, F_taub_cy,
  F_N_a, F_N_b, F_N_cx, F_N_cy, F_visc_int, F_F2, F_beta_a, F_beta_eff_a, F_beta_eff_cx, F_beta_eff_cy,
  F_taudx, F_taudy, F_fg_a, F_fg_cx, F_fg_cy, F_fg_b, F_overb, F_pore, F_Neff, F_front_rhs,
  F_qNW, F_qNE, F_qSW, F_qSE,
  F_xu0, F_xv0, F_xu1, F_xv1, F_u_prev, F_v_prev, F_u_base, F_v_base,
  F_uvav, F_vvav, F_u_vav_a, F_v_vav_a, F_u_surf_a, F_v_surf_a, F_u_base_a, F_v_base_a,
  F_uabs_vav, F_uabs_surf, F_uabs_base, F_R_shear, F_w_surf, F_w_base,
  F_u0, F_v0, F_tbx0, F_tby0,
  F_ru, F_rv, F_rhu, F_rhv, F_pu, F_pv, F_vu, F_vv, F_phu, F_phv, F_shu, F_shv, F_tu, F_tv,
  F_eps_h, F_xbu, F_xbv,
  F2_COUNT
};
enum I2 { I_ice, I_sheet, I_ocean, I_land, I_shelf, I_combi, I_fin, I_fout, I2_COUNT };
enum F3 { G_A3, G_B3, G_visc_a, G_visc_b, G_visc0, G_F1, G_u3, G_v3, G_u3sia, G_v3sia, G_w3, G_u3a, G_v3a, G_u3a0, G_v3a0,
          G_Ti, G_Tpmp, G_Acx, G_Acy, F3_COUNT };

enum Kind { KA, KCX, KCY, KB };

struct HostMember {
  int nx, ny;
};

inline double *f2p(imau_problem *P, int f) { return P->f2.p + (size_t)f * P->na; }
inline int32_t *i2p(imau_problem *P, int f) { return P->i2.p + (size_t)f * P->na; }
inline double *f3p(imau_problem *P, int f) { return P->f3.p + (size_t)f * P->na * P->nz; }

// copy a host Fortran 2D field of kind k into member g's slab (ld ny)
void up2(imau_ctx *c, imau_problem *P, int g, Kind k, const void *src, void *dst_base, size_t esz) {
  if (!src) return;
  const Member &m = P->mem[g];
  int ni = m.nx, nj = m.ny;
  if (k == KCX || k == KB) ni = m.nx - 1;
  if (k == KCY || k == KB) nj = m.ny - 1;
  char *dst = (char *)dst_base + (size_t)m.off * esz;
  HIPCHK(hipMemcpy2DAsync(dst, (size_t)m.ny * esz, src, (size_t)nj * esz, (size_t)nj * esz, ni, hipMemcpyHostToDevice, c->s));
}
void down2(imau_ctx *c, imau_problem *P, int g, Kind k, void *dst, const void *src_base, size_t esz) {
  if (!dst) return;
  if (c->batch) {
    c->batch->push_back(Down2Req{g, (int)k, dst, (const char *)src_base, esz});
    return;
  }
  const Member &m = P->mem[g];
  int ni = m.nx, nj = m.ny;
  if (k == KCX || k == KB) ni = m.nx - 1;
  if (k == KCY || k == KB) nj = m.ny - 1;
  const char *src = (const char *)src_base + (size_t)m.off * esz;
  HIPCHK(hipMemcpy2DAsync(dst, (size_t)nj * esz, src, (size_t)m.ny * esz, (size_t)nj * esz, ni, hipMemcpyDeviceToHost, c->s));
}

// The collected 2D downloads, one device-to-host copy per field: the field's
// whole arena (every member, padding included) goes into pinned memory and the
// host copies each member's part into its array, instead of one pageable
// copy per member and field.  Two pinned buffers let the next field's copy
// run while the host copies out the previous one.
void flush_down2(imau_ctx *c, imau_problem *P, const std::vector<Down2Req> &reqs) {
  std::vector<const char *> srcs;
  std::vector<size_t> eszs;
  for (const Down2Req &r : reqs)
    if (std::find(srcs.begin(), srcs.end(), r.src_base) == srcs.end()) {
      srcs.push_back(r.src_base);
      eszs.push_back(r.esz);
    }
  if (srcs.empty()) return;
  const size_t need = (size_t)P->na * sizeof(double);
  if (c->pin_bytes < need) {
    HIPCHK(hipStreamSynchronize(c->s));
    for (void *&p : c->pin) {
      if (p) HIPCHK(hipHostFree(p));
      p = nullptr;
    }
    c->pin_bytes = 0;
    for (void *&p : c->pin) HIPCHK(hipHostMalloc(&p, need, hipHostMallocDefault));
    c->pin_bytes = need;
  }
  for (hipEvent_t &e : c->pin_ev)
    if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  auto issue = [&](size_t f) {
    HIPCHK(hipMemcpyAsync(c->pin[f & 1], srcs[f], (size_t)P->na * eszs[f], hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipEventRecord(c->pin_ev[f & 1], c->s));
  };
  issue(0);
  for (size_t f = 0; f < srcs.size(); ++f) {
    // buffer (f + 1) & 1 was copied out in the previous pass of this loop
    if (f + 1 < srcs.size()) issue(f + 1);
    HIPCHK(hipEventSynchronize(c->pin_ev[f & 1]));
    const char *h = (const char *)c->pin[f & 1];
    for (const Down2Req &r : reqs) {
      if (r.src_base != srcs[f]) continue;
      const Member &m = P->mem[r.g];
      const int ni = (r.kind == KCX || r.kind == KB) ? m.nx - 1 : m.nx;
      const int nj = (r.kind == KCY || r.kind == KB) ? m.ny - 1 : m.ny;
      const char *src = h + (size_t)m.off * r.esz;
      char *dst = (char *)r.dst;
      if (nj == m.ny)
        std::memcpy(dst, src, (size_t)ni * nj * r.esz);
      else
        for (int i = 0; i < ni; ++i)
          std::memcpy(dst + (size_t)i * nj * r.esz, src + (size_t)i * m.ny * r.esz, (size_t)nj * r.esz);
    }
  }
}
void up3(imau_ctx *c, imau_problem *P, int g, Kind k, const double *src, double *dst) {
  if (!src) return;
  const Member &m = P->mem[g];
  int ni = m.nx, nj = m.ny;
  if (k == KCX) ni = m.nx - 1;
  if (k == KCY) nj = m.ny - 1;
  size_t n = (size_t)ni * nj * P->nz;
  HIPCHK(hipMemcpyAsync(P->stage.p, src, n * sizeof(double), hipMemcpyHostToDevice, c->s));
  imau::h2d_3d(c->s, P->mem_d.p, g, P->stage.p, dst, ni, nj, nj, P->nz);
  HIPCHK(hipGetLastError());
}
void down3(imau_ctx *c, imau_problem *P, int g, Kind k, double *dst, const double *src) {
  if (!dst) return;
  const Member &m = P->mem[g];
  int ni = m.nx, nj = m.ny;
  if (k == KCX) ni = m.nx - 1;
  if (k == KCY) nj = m.ny - 1;
  size_t n = (size_t)ni * nj * P->nz;
  imau::d2h_3d(c->s, P->mem_d.p, g, src, P->stage.p, ni, nj, nj, P->nz);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(dst, P->stage.p, n * sizeof(double), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
}

void check_cfg(const imau_cfg *cfg) {
  if (!cfg) fail(IMAU_ERR_ARG, "cfg is NULL");
  if (cfg->nz < 2 || cfg->nz > IMAU_NZ_MAX) fail(IMAU_ERR_CONFIG, "nz=%d outside [2,%d]", cfg->nz, IMAU_NZ_MAX);
  if (cfg->choice_ice_margin != IMAU_MARGIN_BC && cfg->choice_ice_margin != IMAU_MARGIN_INFINITE_SLAB)
    fail(IMAU_ERR_CONFIG, "unknown choice_ice_margin %d", cfg->choice_ice_margin);
  if (cfg->choice_sliding_law < 0 || cfg->choice_sliding_law > 7)
    fail(IMAU_ERR_CONFIG, "unknown choice_sliding_law %d", cfg->choice_sliding_law);
  if (cfg->choice_sliding_law == IMAU_SLID_IDEALISED &&
      (cfg->choice_idealised_sliding_law < IMAU_IDEAL_ISMIP_HOM_C || cfg->choice_idealised_sliding_law > IMAU_IDEAL_ISMIP_HOM_F))
    fail(IMAU_ERR_CONFIG, "unknown choice_idealised_sliding_law %d", cfg->choice_idealised_sliding_law);
  if (cfg->choice_basal_hydrology < 0 || cfg->choice_basal_hydrology > 2)
    fail(IMAU_ERR_CONFIG, "unknown choice_basal_hydrology %d", cfg->choice_basal_hydrology);
  if (cfg->choice_basal_roughness < 0 || cfg->choice_basal_roughness > 2)
    fail(IMAU_ERR_CONFIG, "unknown choice_basal_roughness %d", cfg->choice_basal_roughness);
  const int law = cfg->choice_sliding_law;
  if (cfg->choice_basal_roughness == IMAU_ROUGH_PARAMETERISED && law != IMAU_SLID_NO_SLIDING && law != IMAU_SLID_IDEALISED) {
    const int pr = cfg->choice_param_basal_roughness;
    if (pr < 0 || pr > 2) fail(IMAU_ERR_CONFIG, "unknown choice_param_basal_roughness %d", pr);
    // src/basal_conditions_and_sliding_module.f90:789-792
    if (pr == IMAU_PROUGH_MARTIN2011 && law != IMAU_SLID_COULOMB && law != IMAU_SLID_COULOMB_REGULARISED &&
        law != IMAU_SLID_ZOET_IVERSON)
      fail(IMAU_ERR_CONFIG, "Martin2011 roughness is only applicable when choice_sliding_law = Coulomb, "
                            "Coulomb_regularised or Zoet-Iverson");
    // :896-898
    if (pr == IMAU_PROUGH_MISMIPPLUS && law != IMAU_SLID_WEERTMAN && law != IMAU_SLID_TSAI2015 &&
        law != IMAU_SLID_SCHOOF2005)
      fail(IMAU_ERR_CONFIG, "MISMIP+ roughness is only defined when choice_sliding_law = Weertman, Tsai2015 or "
                            "Schoof2005");
  }
  const int *bcs = &cfg->DIVA_boundary_BC_u_west;
  for (int q = 0; q < 8; ++q)
    if (bcs[q] < 0 || bcs[q] > 2) fail(IMAU_ERR_CONFIG, "unknown DIVA boundary condition %d", bcs[q]);
  if (cfg->DIVA_choice_matrix_solver != IMAU_SOLVER_SOR && cfg->DIVA_choice_matrix_solver != IMAU_SOLVER_PETSC)
    fail(IMAU_ERR_CONFIG, "unknown DIVA_choice_matrix_solver %d", cfg->DIVA_choice_matrix_solver);
  if (!(cfg->picard_relax_max >= 0.0 && cfg->picard_relax_max <= 1.0))
    fail(IMAU_ERR_CONFIG, "picard_relax_max must lie in [0, 1], got %g", cfg->picard_relax_max);
  if (cfg->picard_relax_max > 0.0 && cfg->stop_rule != IMAU_STOP_RESIDUAL)
    fail(IMAU_ERR_CONFIG, "adaptive Picard relaxation (picard_relax_max > 0) needs stop_rule IMAU_STOP_RESIDUAL");
  if (cfg->mg_rebuild < 1) fail(IMAU_ERR_CONFIG, "mg_rebuild must be >= 1, got %d", cfg->mg_rebuild);
  if (cfg->mg_graph != 0 && cfg->mg_graph != 1) fail(IMAU_ERR_CONFIG, "mg_graph must be 0 or 1, got %d", cfg->mg_graph);
  if (!(cfg->anderson_beta > 0.0 && cfg->anderson_beta <= 1.0))
    fail(IMAU_ERR_CONFIG, "anderson_beta must lie in (0, 1], got %g", cfg->anderson_beta);
  if (!(cfg->anderson_restart >= 1.0))
    fail(IMAU_ERR_CONFIG, "anderson_restart must be >= 1, got %g", cfg->anderson_restart);
  if (cfg->warm_column_iters < 0 || cfg->warm_column_iters > 50)
    fail(IMAU_ERR_CONFIG, "warm_column_iters must lie in [0, 50], got %d", cfg->warm_column_iters);
  if (cfg->warm_visc_iters < 0 || cfg->warm_visc_iters > 50)
    fail(IMAU_ERR_CONFIG, "warm_visc_iters must lie in [0, 50], got %d", cfg->warm_visc_iters);
  if (cfg->anderson_m < 0 || cfg->anderson_m > IMAU_AND_SLOTS - 1)
    fail(IMAU_ERR_CONFIG, "anderson_m must lie in [0, %d], got %d", IMAU_AND_SLOTS - 1, cfg->anderson_m);
  if (cfg->anderson_m > 0 && cfg->stop_rule != IMAU_STOP_RESIDUAL)
    fail(IMAU_ERR_CONFIG, "Anderson acceleration (anderson_m > 0) needs stop_rule IMAU_STOP_RESIDUAL");
  if (cfg->anderson_m > 0 && cfg->picard_relax_max > 0.0)
    fail(IMAU_ERR_CONFIG, "anderson_m and picard_relax_max exclude each other");
  if (cfg->precision < IMAU_PREC_FP64 || cfg->precision > IMAU_PREC_MIXED_VISC)
    fail(IMAU_ERR_CONFIG, "unknown precision %d", cfg->precision);
  if (cfg->ksp_type != 0 && cfg->ksp_type != 1)
    fail(IMAU_ERR_CONFIG, "ksp_type must be 0 (BiCGStab) or 1 (GCR), got %d", cfg->ksp_type);
  if (cfg->mg_smoother != 0 && cfg->mg_smoother != 1)
    fail(IMAU_ERR_CONFIG, "mg_smoother must be 0 (fused sweep) or 1 (one launch per colour), got %d", cfg->mg_smoother);
  if (cfg->stall_window < 0 || cfg->stall_window > 1000)
    fail(IMAU_ERR_CONFIG, "stall_window must lie in [0, 1000], got %d", cfg->stall_window);
  if (cfg->stall_window > 0 && !(cfg->stall_drop > 1.0))
    fail(IMAU_ERR_CONFIG, "stall_drop must be > 1, got %g", cfg->stall_drop);
  if (!(cfg->stall_relax >= 0.0 && cfg->stall_relax <= 1.0))
    fail(IMAU_ERR_CONFIG, "stall_relax must lie in [0, 1], got %g", cfg->stall_relax);
  if (cfg->mg_levels != 0 && cfg->mg_levels != 1)
    fail(IMAU_ERR_CONFIG, "mg_levels must be 0 (multigrid, automatic depth) or 1 (no multigrid), got %d",
         cfg->mg_levels);
  for (int k = 1; k < cfg->nz; ++k)
    if (!(cfg->zeta[k] > cfg->zeta[k - 1])) fail(IMAU_ERR_CONFIG, "zeta must increase");
}

KP make_kp(const imau_cfg *c) {
  KP k{};
  k.nz = c->nz;
  for (int q = 0; q < c->nz; ++q) k.zeta[q] = c->zeta[q];
  k.dyn = c->choice_ice_dynamics;
  k.margin = c->choice_ice_margin;
  k.slid = c->choice_sliding_law;
  k.ideal = c->choice_idealised_sliding_law;
  k.hydro = c->choice_basal_hydrology;
  k.rough = c->choice_basal_roughness;
  k.prough = c->choice_param_basal_roughness;
  k.bc_u[0] = c->DIVA_boundary_BC_u_west; k.bc_u[1] = c->DIVA_boundary_BC_u_east;
  k.bc_u[2] = c->DIVA_boundary_BC_u_south; k.bc_u[3] = c->DIVA_boundary_BC_u_north;
  k.bc_v[0] = c->DIVA_boundary_BC_v_west; k.bc_v[1] = c->DIVA_boundary_BC_v_east;
  k.bc_v[2] = c->DIVA_boundary_BC_v_south; k.bc_v[3] = c->DIVA_boundary_BC_v_north;
  k.cross = c->include_SSADIVA_crossterms;
  k.glsub = c->do_GL_subgrid_friction;
  k.freeze = c->include_basal_freezing;
  k.n_flow = c->n_flow;
  k.eps0 = c->DIVA_epsilon_sq_0;
  k.visc_min = c->DIVA_visc_eff_min;
  k.beta_max = c->DIVA_beta_max;
  k.vel_max = c->DIVA_vel_max;
  k.relax = c->DIVA_visc_it_relax;
  k.delta_v = c->slid_delta_v;
  k.wm = c->slid_Weertman_m;
  k.q_plastic = c->slid_Coulomb_reg_q_plastic;
  k.u_threshold = c->slid_Coulomb_reg_u_threshold;
  k.zi_ut = c->slid_ZI_ut;
  k.zi_p = c->slid_ZI_p;
  k.dT_freeze = c->deltaT_basal_freezing;
  k.hydro_min = c->Martin2011_hydro_Hb_min;
  k.hydro_max = c->Martin2011_hydro_Hb_max;
  k.phi_hb_min = c->Martin2011till_phi_Hb_min;
  k.phi_hb_max = c->Martin2011till_phi_Hb_max;
  k.phi_min = c->Martin2011till_phi_min;
  k.phi_max = c->Martin2011till_phi_max;
  k.ssa_L = c->SSA_icestream_L;
  k.ssa_m = c->SSA_icestream_m;
  k.ismip_L = c->ISMIP_HOM_L;
  // F_int_min, src/ice_velocity_module.f90:980
  double f[IMAU_NZ_MAX];
  for (int q = 0; q < c->nz; ++q) f[q] = (1.0 / c->DIVA_visc_eff_min) * (c->zeta[q] * c->zeta[q]);
  k.F_int_min = vint(c->nz, c->zeta, f);
  return k;
}

void wire(imau_problem *P) {
  Fields &F = P->F;
  F.Hi = f2p(P, F_Hi); F.Hb = f2p(P, F_Hb); F.SL = f2p(P, F_SL); F.Hs = f2p(P, F_Hs); F.TAF = f2p(P, F_TAF);
  F.dHs_dx_a = f2p(P, F_dHs_dx_a); F.dHs_dy_a = f2p(P, F_dHs_dy_a);
  F.dHi_dx_a = f2p(P, F_dHi_dx_a); F.dHi_dy_a = f2p(P, F_dHi_dy_a);
  F.dHb_dt = f2p(P, F_dHb_dt); F.A_vav = f2p(P, F_A_vav);
  F.m_ice = i2p(P, I_ice); F.m_sheet = i2p(P, I_sheet); F.m_ocean = i2p(P, I_ocean);
  F.Hi_cx = f2p(P, F_Hi_cx); F.dHs_dx_cx = f2p(P, F_dHs_dx_cx); F.dHs_dy_cx = f2p(P, F_dHs_dy_cx);
  F.Hi_cy = f2p(P, F_Hi_cy); F.dHs_dx_cy = f2p(P, F_dHs_dx_cy); F.dHs_dy_cy = f2p(P, F_dHs_dy_cy);
  F.Hi_b = f2p(P, F_Hi_b);
  F.Ti3 = f3p(P, G_Ti); F.Tpmp3 = f3p(P, G_Tpmp);
  F.A3 = f3p(P, G_A3);
  F.B3 = f3p(P, G_B3);
  F.beta_sq = f2p(P, F_beta_sq); F.alpha_sq = f2p(P, F_alpha_sq); F.phi_fric = f2p(P, F_phi_fric); F.tauc = f2p(P, F_tauc);
  F.u = f2p(P, F_u); F.v = f2p(P, F_v); F.taub_cx = f2p(P, F_taub_cx); F.taub_cy = f2p(P, F_taub_cy);
  F.visc_a = f3p(P, G_visc_a); F.visc_b = f3p(P, G_visc_b);
  F.N_a = f2p(P, F_N_a); F.N_b = f2p(P, F_N_b); F.N_cx = f2p(P, F_N_cx); F.N_cy = f2p(P, F_N_cy);
  F.eps_h = f2p(P, F_eps_h);
  F.visc_int = f2p(P, F_visc_int); F.F2 = f2p(P, F_F2); F.beta_a = f2p(P, F_beta_a); F.beta_eff_a = f2p(P, F_beta_eff_a);
  F.beta_eff_cx = f2p(P, F_beta_eff_cx); F.beta_eff_cy = f2p(P, F_beta_eff_cy);
  F.taudx = f2p(P, F_taudx); F.taudy = f2p(P, F_taudy);
  F.fg_a = f2p(P, F_fg_a); F.fg_cx = f2p(P, F_fg_cx); F.fg_cy = f2p(P, F_fg_cy); F.fg_b = f2p(P, F_fg_b);
  F.overb = f2p(P, F_overb); F.pore = f2p(P, F_pore); F.Neff = f2p(P, F_Neff); F.front_rhs = f2p(P, F_front_rhs);
  F.qNW = f2p(P, F_qNW); F.qNE = f2p(P, F_qNE); F.qSW = f2p(P, F_qSW); F.qSE = f2p(P, F_qSE);
  F.combi = i2p(P, I_combi); F.fin = i2p(P, I_fin); F.fout = i2p(P, I_fout);
  F.rt_u = P->rt.p; F.rt_v = P->rt.p + P->na;
  F.xu0 = f2p(P, F_xu0); F.xv0 = f2p(P, F_xv0); F.xu1 = f2p(P, F_xu1); F.xv1 = f2p(P, F_xv1);
  F.u_prev = f2p(P, F_u_prev); F.v_prev = f2p(P, F_v_prev); F.u_base = f2p(P, F_u_base); F.v_base = f2p(P, F_v_base);
  OutF &O = P->O;
  O.F1 = f3p(P, G_F1); O.u3 = f3p(P, G_u3); O.v3 = f3p(P, G_v3); O.u3sia = f3p(P, G_u3sia); O.v3sia = f3p(P, G_v3sia);
  O.w3 = f3p(P, G_w3); O.u3a = f3p(P, G_u3a); O.v3a = f3p(P, G_v3a);
  O.uvav = f2p(P, F_uvav); O.vvav = f2p(P, F_vvav);
  O.u_vav_a = f2p(P, F_u_vav_a); O.v_vav_a = f2p(P, F_v_vav_a); O.u_surf_a = f2p(P, F_u_surf_a);
  O.v_surf_a = f2p(P, F_v_surf_a); O.u_base_a = f2p(P, F_u_base_a); O.v_base_a = f2p(P, F_v_base_a);
  O.uabs_vav = f2p(P, F_uabs_vav); O.uabs_surf = f2p(P, F_uabs_surf); O.uabs_base = f2p(P, F_uabs_base);
  O.R_shear = f2p(P, F_R_shear); O.w_surf = f2p(P, F_w_surf); O.w_base = f2p(P, F_w_base);
  P->Acx3 = f3p(P, G_Acx); P->Acy3 = f3p(P, G_Acy);
  P->u0 = f2p(P, F_u0); P->v0 = f2p(P, F_v0); P->tbx0 = f2p(P, F_tbx0); P->tby0 = f2p(P, F_tby0);
  P->visc0 = f3p(P, G_visc0); P->u3a0 = f3p(P, G_u3a0); P->v3a0 = f3p(P, G_v3a0);
}

bool uses_ssa_field(int dyn) { return dyn == IMAU_DYN_SSA || dyn == IMAU_DYN_SIA_SSA; }

// buffers that only some configurations use (GCR directions, Anderson
// slots, the FP32 viscosity), allocated before a solve's first kernel since
// DBuf::alloc synchronises the device: at create for the create-time
// configuration, and at the start of a solve whose configuration needs more
void reserve_buffers(imau_problem *P, const imau_cfg *cfg) {
  const long long na = P->na, n3 = na * (long long)P->nz;
  if (cfg->DIVA_choice_matrix_solver != IMAU_SOLVER_SOR && cfg->ksp_type == 1 && !P->gcr_dirs.p)
    P->gcr_dirs.alloc((size_t)4 * IMAU_GCR_M * na);
  if (cfg->stop_rule == IMAU_STOP_RESIDUAL && cfg->anderson_m > 0) {
    if (!P->and_slots.p) P->and_slots.alloc((size_t)4 * IMAU_AND_SLOTS * na);
    if (!P->and_w.p) P->and_w.alloc((size_t)2 * IMAU_AND_SLOTS * P->ngrids);
  }
  if (cfg->precision == IMAU_PREC_MIXED_VISC && !P->visc32.p) P->visc32.alloc((size_t)3 * n3);
}

// ISMIP-HOM E reads its slip columns from the 51-line Arolla file
// (src/basal_conditions_and_sliding_module.f90:1372-1389)
void check_ismip_e(const imau_problem *P, const imau_cfg *cfg) {
  if (cfg->choice_sliding_law != IMAU_SLID_IDEALISED || cfg->choice_idealised_sliding_law != IMAU_IDEAL_ISMIP_HOM_E)
    return;
  for (int g = 0; g < P->ngrids; ++g) {
    if (!P->has_slides[g])
      fail(IMAU_ERR_ARG, "member %d: ISMIP_HOM_E needs ISMIP_HOM_E_slides at create", g);
    if (P->mem[g].nx != 51)
      fail(IMAU_ERR_CONFIG, "member %d: ISMIP_HOM_E: the Arolla file has 51 lines, one per grid column (nx=%d)", g,
           P->mem[g].nx);
  }
}

void upload(imau_ctx *c, imau_problem *P, const imau_cfg *cfg, const imau_ice_io *ios) {
  const int nz = P->nz;
  for (int g = 0; g < P->ngrids; ++g) {
    const imau_ice_io &io = ios[g];
    if (!io.Hi_a || !io.Hb_a || !io.SL_a || !io.Hs_a || !io.TAF_a || !io.mask_ice_a || !io.mask_sheet_a ||
        !io.mask_ocean_a || !io.Hi_cx || !io.Hi_cy || !io.Hi_b || !io.dHs_dx_cx || !io.dHs_dy_cy || !io.dHs_dx_cy ||
        !io.dHs_dy_cx || !io.dHs_dx_a || !io.dHs_dy_a || !io.dHi_dx_a || !io.dHi_dy_a)
      fail(IMAU_ERR_ARG, "member %d: a required geometry input is NULL", g);
    if ((cfg->choice_ice_dynamics == IMAU_DYN_DIVA || uses_ssa_field(cfg->choice_ice_dynamics)) && !io.A_flow_3D_a)
      fail(IMAU_ERR_ARG, "member %d: A_flow_3D_a is NULL", g);
    if (cfg->include_basal_freezing && (!io.Ti_a || !io.Ti_pmp_a))
      fail(IMAU_ERR_ARG, "member %d: include_basal_freezing needs Ti_a and Ti_pmp_a", g);
    const size_t d = sizeof(double), q = sizeof(int32_t);
    up2(c, P, g, KA, io.Hi_a, f2p(P, F_Hi), d);
    up2(c, P, g, KA, io.Hb_a, f2p(P, F_Hb), d);
    up2(c, P, g, KA, io.SL_a, f2p(P, F_SL), d);
    up2(c, P, g, KA, io.Hs_a, f2p(P, F_Hs), d);
    up2(c, P, g, KA, io.TAF_a, f2p(P, F_TAF), d);
    up2(c, P, g, KA, io.dHs_dx_a, f2p(P, F_dHs_dx_a), d);
    up2(c, P, g, KA, io.dHs_dy_a, f2p(P, F_dHs_dy_a), d);
    up2(c, P, g, KA, io.dHi_dx_a, f2p(P, F_dHi_dx_a), d);
    up2(c, P, g, KA, io.dHi_dy_a, f2p(P, F_dHi_dy_a), d);
    up2(c, P, g, KA, io.dHb_dt_a, f2p(P, F_dHb_dt), d);
    up2(c, P, g, KA, io.A_flow_vav_a, f2p(P, F_A_vav), d);
    up2(c, P, g, KA, io.mask_ice_a, i2p(P, I_ice), q);
    up2(c, P, g, KA, io.mask_sheet_a, i2p(P, I_sheet), q);
    up2(c, P, g, KA, io.mask_ocean_a, i2p(P, I_ocean), q);
    up2(c, P, g, KCX, io.Hi_cx, f2p(P, F_Hi_cx), d);
    up2(c, P, g, KCX, io.dHs_dx_cx, f2p(P, F_dHs_dx_cx), d);
    up2(c, P, g, KCX, io.dHs_dy_cx, f2p(P, F_dHs_dy_cx), d);
    up2(c, P, g, KCY, io.Hi_cy, f2p(P, F_Hi_cy), d);
    up2(c, P, g, KCY, io.dHs_dx_cy, f2p(P, F_dHs_dx_cy), d);
    up2(c, P, g, KCY, io.dHs_dy_cy, f2p(P, F_dHs_dy_cy), d);
    up2(c, P, g, KB, io.Hi_b, f2p(P, F_Hi_b), d);
    up2(c, P, g, KA, io.beta_sq_a, f2p(P, F_beta_sq), d);
    up2(c, P, g, KA, io.alpha_sq_a, f2p(P, F_alpha_sq), d);
    up2(c, P, g, KA, io.phi_fric_a, f2p(P, F_phi_fric), d);
    up2(c, P, g, KA, io.tauc_a, f2p(P, F_tauc), d);
    const bool ssa = uses_ssa_field(cfg->choice_ice_dynamics);
    up2(c, P, g, KCX, ssa ? io.u_SSA_cx : io.u_vav_cx, P->u0, d);
    up2(c, P, g, KCY, ssa ? io.v_SSA_cy : io.v_vav_cy, P->v0, d);
    if (ssa && P->err.p) {  // ice%u_vav_cx of the state: read by the SSA error estimate (:1626)
      up2(c, P, g, KCX, io.u_vav_cx, P->O.uvav, d);
      up2(c, P, g, KCY, io.v_vav_cy, P->O.vvav, d);
    }
    up2(c, P, g, KCX, io.taub_cx, P->tbx0, d);
    up2(c, P, g, KCY, io.taub_cy, P->tby0, d);
    up2(c, P, g, KA, io.N_a, f2p(P, F_N_a), d);              // BC margin keeps stale N_a of ice-free cells
    up2(c, P, g, KA, io.visc_eff_int_a, f2p(P, F_visc_int), d);
    up3(c, P, g, KA, io.A_flow_3D_a, f3p(P, G_A3));
    up3(c, P, g, KA, io.visc_eff_3D_a, P->visc0);
    up3(c, P, g, KA, io.u_3D_a, P->u3a0);                  // previous a-grid velocities (read by w_3D)
    up3(c, P, g, KA, io.v_3D_a, P->v3a0);
    up3(c, P, g, KCX, io.A_flow_3D_cx, P->Acx3);
    up3(c, P, g, KCY, io.A_flow_3D_cy, P->Acy3);
    if (cfg->choice_ice_dynamics == IMAU_DYN_SIA_SSA) {  // SIA columns feed the hybrid SSA (:444-451)
      up3(c, P, g, KCX, io.u_3D_SIA_cx, f3p(P, G_u3sia));
      up3(c, P, g, KCY, io.v_3D_SIA_cy, f3p(P, G_v3sia));
    }
    if (g == 0) {
      P->sor_nit_ = io.DIVA_SOR_nit > 0 ? io.DIVA_SOR_nit : 10000;
      P->sor_tol_ = io.DIVA_SOR_tol;
      P->sor_omega_ = io.DIVA_SOR_omega > 0.0 ? io.DIVA_SOR_omega : 1.3;
      P->petsc_rtol_ = io.DIVA_PETSc_rtol;
      P->petsc_abstol_ = io.DIVA_PETSc_abstol;
    }
    if (cfg->include_basal_freezing) {
      up3(c, P, g, KA, io.Ti_a, f3p(P, G_Ti));
      up3(c, P, g, KA, io.Ti_pmp_a, f3p(P, G_Tpmp));
    }
    // trivial members (no grounded ice): src/ice_velocity_module.f90:276-285
    long long s = 0;
    const Member &m = P->mem[g];
    for (long long p = 0; p < (long long)m.nx * m.ny; ++p) s += io.mask_sheet_a[p];
    P->sheet_any[g] = s > 0;
  }
  (void)nz;
  HIPCHK(hipStreamSynchronize(c->s));
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
#define ABI_TRY try {
#define ABI_CATCH(ctx)                                        \
  }                                                           \
  catch (const Err &e) {                                      \
    if (ctx) (ctx)->err = e.msg; else g_init_err = e.msg;     \
    return e.code;                                            \
  }                                                           \
  catch (const std::exception &e) {                           \
    if (ctx) (ctx)->err = e.what(); else g_init_err = e.what(); \
    return IMAU_ERR_DEVICE;                                   \
  }

extern "C" {

void imau_vel_default_config(imau_cfg *c) {
  static const double zeta[15] = {0.00, 0.10, 0.20, 0.30, 0.40, 0.50, 0.60, 0.70, 0.80, 0.90, 0.925, 0.95, 0.975, 0.99, 1.00};
  std::memset(c, 0, sizeof *c);
  c->choice_ice_dynamics = IMAU_DYN_DIVA;
  c->choice_ice_margin = IMAU_MARGIN_INFINITE_SLAB;
  c->choice_sliding_law = IMAU_SLID_COULOMB_REGULARISED;
  c->choice_basal_hydrology = IMAU_HYDRO_MARTIN2011;
  c->choice_basal_roughness = IMAU_ROUGH_PARAMETERISED;
  c->choice_param_basal_roughness = IMAU_PROUGH_MARTIN2011;
  c->DIVA_choice_matrix_solver = IMAU_SOLVER_PETSC;
  c->include_SSADIVA_crossterms = 1;
  c->do_GL_subgrid_friction = 1;
  c->include_basal_freezing = 1;
  c->nz = 15;
  for (int k = 0; k < 15; ++k) c->zeta[k] = zeta[k];
  c->n_flow = 3.0;
  c->DIVA_visc_it_norm_dUV_tol = 1e-2;
  c->DIVA_visc_it_nit = 50;
  c->DIVA_visc_it_relax = 0.4;
  c->DIVA_epsilon_sq_0 = 1e-15;
  c->DIVA_visc_eff_min = 1e3;
  c->DIVA_beta_max = 1e20;
  c->DIVA_vel_max = 5000.0;
  c->slid_delta_v = 1.0e-3;
  c->slid_Weertman_m = 3.0;
  c->slid_Coulomb_reg_q_plastic = 0.3;
  c->slid_Coulomb_reg_u_threshold = 100.0;
  c->slid_ZI_ut = 200.0;
  c->slid_ZI_p = 5.0;
  c->deltaT_basal_freezing = 2.0;
  c->Martin2011_hydro_Hb_min = 0.0;
  c->Martin2011_hydro_Hb_max = 1000.0;
  c->Martin2011till_phi_Hb_min = -1000.0;
  c->Martin2011till_phi_Hb_max = 0.0;
  c->Martin2011till_phi_min = 5.0;
  c->Martin2011till_phi_max = 20.0;
  c->SSA_icestream_L = 100000.0;
  c->SSA_icestream_m = 0.5;
  c->ISMIP_HOM_L = 160000.0;
  c->precision = IMAU_PREC_FP64;
  c->stop_rule = IMAU_STOP_PICARD;
  c->nonlinear_rtol = 1e-4;
  c->picard_relax_max = 0.0;
  c->mg_smoother = 0;
  c->ksp_type = 0;
  c->anderson_m = 0;
  c->mg_rebuild = 1;
  c->mg_graph = 1;
  c->warm_visc_iters = 0;
  c->anderson_beta = 1.0;
  c->anderson_restart = 1.5;
  c->warm_column_iters = 0;
  c->stall_window = 0;
  c->stall_drop = 10.0;
  c->stall_relax = 0.0;
}

int imau_vel_init(const imau_grid *grid, const imau_cfg *cfg, imau_ctx **out) {
  imau_ctx *ctx = nullptr;
  ABI_TRY
  if (!out) fail(IMAU_ERR_ARG, "ctx output pointer is NULL");
  *out = nullptr;
  if (grid && (grid->nx < 4 || grid->ny < 4)) fail(IMAU_ERR_ARG, "grid must be at least 4x4 (got %dx%d)", grid->nx, grid->ny);
  if (cfg) check_cfg(cfg);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) fail(IMAU_ERR_NO_DEVICE, "no HIP device visible");
  ctx = new imau_ctx();
  HIPCHK(hipGetDevice(&ctx->device));
  HIPCHK(hipStreamCreateWithFlags(&ctx->s, hipStreamNonBlocking));
  ctx->scratch = scratch_address(ctx->device);
  *out = ctx;
  return IMAU_OK;
  }
  catch (const Err &e) {
    g_init_err = e.msg;
    delete ctx;
    return e.code;
  }
}

void imau_vel_finalize(imau_ctx *ctx) {
  if (!ctx) return;
  if (ctx->s) (void)hipStreamSynchronize(ctx->s);
  for (hipEvent_t e : ctx->ev) (void)hipEventDestroy(e);
  for (hipEvent_t e : ctx->pin_ev)
    if (e) (void)hipEventDestroy(e);
  for (void *p : ctx->pin)
    if (p) (void)hipHostFree(p);
  imau_vel_problem_destroy(ctx->cached);
  if (ctx->s) (void)hipStreamDestroy(ctx->s);
  delete ctx;
}

const char *imau_vel_last_error(const imau_ctx *ctx) { return ctx ? ctx->err.c_str() : g_init_err.c_str(); }

int imau_vel_kernel_timing(imau_ctx *ctx, int32_t what, int64_t *launches, double *ms, int64_t *cells) {
  if (!ctx) return IMAU_ERR_ARG;
  if (what < IMAU_TIME_OFF || what > IMAU_TIME_VISC) return IMAU_ERR_ARG;
  if (launches) *launches = ctx->sweep_launches;
  if (ms) *ms = ctx->sweep_ms;
  if (cells) *cells = ctx->sweep_cells;
  ctx->ktime = what;
  ctx->sweep_launches = 0;
  ctx->sweep_ms = 0.0;
  ctx->sweep_cells = 0;
  return IMAU_OK;
}

int imau_vel_problem_create(imau_ctx *ctx, int32_t ngrids, const imau_grid *grids, const imau_cfg *cfg,
                            const imau_ice_io *ios, imau_problem **prob) {
  imau_problem *P = nullptr;
  ABI_TRY
  if (!ctx || !grids || !ios || !prob) fail(IMAU_ERR_ARG, "NULL argument");
  if (ngrids < 1) fail(IMAU_ERR_ARG, "ngrids=%d", ngrids);
  check_cfg(cfg);
  P = new imau_problem();
  P->ngrids = ngrids;
  P->nz = cfg->nz;
  long long off = 0, offx = 0, offy = 0;
  std::vector<double> xs, ys;
  size_t max3 = 0;
  for (int g = 0; g < ngrids; ++g) {
    const imau_grid &gr = grids[g];
    if (gr.nx < 4 || gr.ny < 4) fail(IMAU_ERR_ARG, "member %d: grid must be at least 4x4 (got %dx%d)", g, gr.nx, gr.ny);
    if (!(gr.dx > 0.0)) fail(IMAU_ERR_ARG, "member %d: dx must be positive", g);
    if (!gr.x || !gr.y) fail(IMAU_ERR_ARG, "member %d: grid coordinates are NULL", g);
    Member m{};
    m.nx = gr.nx; m.ny = gr.ny; m.dx = gr.dx;
    m.off = off; m.off3 = off * cfg->nz; m.offx = offx; m.offy = offy;
    P->mem.push_back(m);
    P->tstart.push_back((int)P->tiles.size());
    for (int i0 = 0; i0 < gr.nx; i0 += IMAU_TI)
      for (int j0 = 0; j0 < gr.ny; j0 += IMAU_TJ) P->tiles.push_back(Tile{g, i0, j0, 0});
    // pad each slab to a multiple of 64 doubles (512 B) for aligned member starts
    long long na = (long long)gr.nx * gr.ny;
    off += (na + 63) / 64 * 64;
    xs.insert(xs.end(), gr.x, gr.x + gr.nx);
    ys.insert(ys.end(), gr.y, gr.y + gr.ny);
    offx += gr.nx; offy += gr.ny;
    max3 = std::max(max3, (size_t)na * cfg->nz);
  }
  P->tstart.push_back((int)P->tiles.size());
  P->na = off;
  P->sheet_any.assign(ngrids, 0);
  P->picard_its.assign(ngrids, 0);
  P->restarts.assign(ngrids, 0);
  P->sweeps.assign(ngrids, 0);
  P->resid_uv.assign(ngrids, 0.0);
  P->nl_resid.assign(ngrids, NAN);
  P->mem_d.alloc(ngrids);
  P->tiles_d.alloc(P->tiles.size());
  P->tstart_d.alloc(P->tstart.size());
  P->xg.alloc(xs.size());
  {
    std::vector<double> eb(xs.size(), 1e30);
    P->has_slides.assign(ngrids, 0);
    for (int g = 0; g < ngrids; ++g)
      if (const int32_t *sl = ios[g].ISMIP_HOM_E_slides) {
        P->has_slides[g] = 1;
        for (int i = 0; i < P->mem[g].nx; ++i) eb[P->mem[g].offx + i] = sl[i] == 1 ? 0.0 : 1e30;
      }
    P->ismip_e.alloc(eb.size());
    HIPCHK(hipMemcpy(P->ismip_e.p, eb.data(), eb.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  P->yg.alloc(ys.size());
  P->f2.alloc((size_t)F2_COUNT * P->na);
  P->i2.alloc((size_t)I2_COUNT * P->na);
  P->rt.alloc((size_t)2 * P->na);
  P->f3.alloc((size_t)F3_COUNT * P->na * cfg->nz);
  P->stage.alloc(max3);
  P->partial.alloc((size_t)(2 * IMAU_GCR_M + 2) * P->tiles.size());
  P->msum.alloc((size_t)4 * ngrids);
  P->resmax.alloc(ngrids);
  P->act_pic.alloc(ngrids);
  P->act_lin.alloc(ngrids);
  P->sc.alloc((size_t)8 * ngrids);
  P->relax_d.alloc(ngrids);
  P->restart_d.alloc(ngrids);
  P->stop_d.alloc(ngrids);
  P->lin_d.alloc(ngrids);
  HIPCHK(hipMemcpy(P->mem_d.p, P->mem.data(), ngrids * sizeof(Member), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(P->tiles_d.p, P->tiles.data(), P->tiles.size() * sizeof(Tile), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(P->tstart_d.p, P->tstart.data(), P->tstart.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(P->xg.p, xs.data(), xs.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(P->yg.p, ys.data(), ys.size() * sizeof(double), hipMemcpyHostToDevice));
  build_mg(P);
  wire(P);
  P->F.ismip_e = P->ismip_e.p;
  // estimate_visc_iter_UV_errors is computed when some member asks for it
  bool want_err = false;
  for (int g = 0; g < ngrids; ++g) want_err = want_err || ios[g].DIVA_err_cx || ios[g].DIVA_err_cy;
  if (want_err) {
    P->err.alloc((size_t)4 * P->na);
    P->err_slot.assign(ngrids, 0);
    P->F.chk_u = P->O.uvav;
    P->F.chk_v = P->O.vvav;
  }
  reserve_buffers(P, cfg);
  upload(ctx, P, cfg, ios);
  if (want_err)  // slot 0 starts with the caller's values (kept by a member that never enters its loop)
    for (int g = 0; g < ngrids; ++g) {
      up2(ctx, P, g, KCX, ios[g].DIVA_err_cx, P->err.p, sizeof(double));
      up2(ctx, P, g, KCY, ios[g].DIVA_err_cy, P->err.p + P->na, sizeof(double));
    }
  HIPCHK(hipStreamSynchronize(ctx->s));  // the caller's arrays may go once create returns
  P->dyn_uploaded = cfg->choice_ice_dynamics;
  P->has_ti = cfg->include_basal_freezing != 0;
  *prob = P;
  return imau_vel_problem_reset_state(ctx, P);
  }
  catch (const Err &e) {
    if (ctx) ctx->err = e.msg;
    delete P;
    return e.code;
  }
}

void imau_vel_problem_destroy(imau_problem *P) { delete P; }

int imau_vel_problem_restarts(imau_ctx *ctx, const imau_problem *P, int32_t *restarts) {
  if (!ctx || !P || !restarts) return IMAU_ERR_ARG;
  for (int g = 0; g < P->ngrids; ++g) restarts[g] = P->restarts.empty() ? 0 : P->restarts[g];
  return IMAU_OK;
}

int imau_vel_problem_stats(imau_ctx *ctx, const imau_problem *P, int32_t *picard_iterations,
                           double *nonlinear_residual) {
  ABI_TRY
  if (!ctx || !P) fail(IMAU_ERR_ARG, "NULL argument");
  for (int g = 0; g < P->ngrids; ++g) {
    if (picard_iterations) picard_iterations[g] = P->picard_its.empty() ? 0 : P->picard_its[g];
    if (nonlinear_residual) nonlinear_residual[g] = P->nl_resid.empty() ? NAN : P->nl_resid[g];
  }
  return IMAU_OK;
  ABI_CATCH(ctx)
}

int imau_vel_problem_reset_state(imau_ctx *ctx, imau_problem *P) {
  ABI_TRY
  if (!ctx || !P) fail(IMAU_ERR_ARG, "NULL argument");
  const size_t b2 = (size_t)P->na * sizeof(double), b3 = b2 * P->nz;
  HIPCHK(hipMemcpyAsync(P->F.u, P->u0, b2, hipMemcpyDeviceToDevice, ctx->s));
  HIPCHK(hipMemcpyAsync(P->F.v, P->v0, b2, hipMemcpyDeviceToDevice, ctx->s));
  HIPCHK(hipMemcpyAsync(P->F.taub_cx, P->tbx0, b2, hipMemcpyDeviceToDevice, ctx->s));
  HIPCHK(hipMemcpyAsync(P->F.taub_cy, P->tby0, b2, hipMemcpyDeviceToDevice, ctx->s));
  HIPCHK(hipMemcpyAsync(P->F.visc_a, P->visc0, b3, hipMemcpyDeviceToDevice, ctx->s));
  HIPCHK(hipMemcpyAsync(P->O.u3a, P->u3a0, b3, hipMemcpyDeviceToDevice, ctx->s));
  HIPCHK(hipMemcpyAsync(P->O.v3a, P->v3a0, b3, hipMemcpyDeviceToDevice, ctx->s));
  P->visc_cur = P->F.visc_a;
  P->visc_oth = P->F.visc_b;
  return IMAU_OK;
  ABI_CATCH(ctx)
}

}  // extern "C"

// ---------------------------------------------------------------------------
// the solves
// ---------------------------------------------------------------------------
namespace {

struct Solver {
  imau_ctx *c;
  imau_problem *P;
  const imau_cfg *cfg;
  KP kp;
  imau::Launch L;
  std::vector<int> act_pic, act_lin;
  std::vector<double> host4;

  Solver(imau_ctx *c_, imau_problem *P_, const imau_cfg *cfg_) : c(c_), P(P_), cfg(cfg_) {
    kp = make_kp(cfg);
    L = imau::Launch{c->s, (int)P->tiles.size(), P->mem_d.p, P->tiles_d.p};
    host4.resize(4 * (size_t)P->ngrids);
  }
  imau::Coef coef() const {
    imau::Coef k;
    k.Na = P->F.N_a; k.Nb = P->F.N_b; k.Ncx = P->F.N_cx; k.Ncy = P->F.N_cy;
    k.bcx = P->F.beta_eff_cx; k.bcy = P->F.beta_eff_cy; k.taudx = P->F.taudx; k.taudy = P->F.taudy;
    k.frhs = P->F.front_rhs; k.rt_u = P->F.rt_u; k.rt_v = P->F.rt_v;
    return k;
  }
  // device copies of the activity flags are refreshed only when they change;
  // the staging vectors must outlive the asynchronous copy, hence members
  std::vector<int> last_pic_, last_lin_, stage_pic_, stage_lin_;
  void push_active(int *dev, const std::vector<int> &h) {
    const bool pic = dev == P->act_pic.p;
    std::vector<int> &last = pic ? last_pic_ : last_lin_;
    std::vector<int> &stage = pic ? stage_pic_ : stage_lin_;
    if (last == h) return;
    HIPCHK(hipStreamSynchronize(c->s));   // the previous copy from `stage` has completed
    stage = h;
    last = h;
    HIPCHK(hipMemcpyAsync(dev, stage.data(), stage.size() * sizeof(int), hipMemcpyHostToDevice, c->s));
  }
  void kcheck() { HIPCHK(hipGetLastError()); }
  imau::Launch LV() const { return imau::Launch{c->s, P->nvtiles, P->mem_d.p, P->vtiles.p, P->nmtiles, P->mtiles.p}; }

  std::vector<double> r_abs, b_abs;  // last residual evaluation: ||r||, ||b|| per member

  // per-member relative FP64 residual at the current coefficients and state
  // With GCR in KSP mode the same pass also prepares the linear solve's start
  // vector and residual (lin_r0_), which gcr_solve then takes over.
  bool r0_ready_ = false;
  std::vector<double> lin_r0_;
  void residual(std::vector<double> &r) {
    const bool fuse = cfg->DIVA_choice_matrix_solver != IMAU_SOLVER_SOR && cfg->ksp_type == 1;
    const int NV = fuse ? 3 : 2;
    if (fuse) {
      imau::resid_init(L, kp, coef(), P->F.u, P->F.v, P->F.xu0, P->F.xv0, f2p(P, F_ru), f2p(P, F_rv), P->act_pic.p,
                       P->partial.p);
      imau::member_sum3(c->s, P->ngrids, P->tstart_d.p, P->partial.p, P->msum.p);
    } else {
      imau::resid(L, kp, coef(), P->F.u, P->F.v, P->partial.p);
      imau::member_sum2(c->s, P->ngrids, P->tstart_d.p, P->partial.p, P->msum.p);
    }
    kcheck();
    HIPCHK(hipMemcpyAsync(host4.data(), P->msum.p, NV * P->ngrids * sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    r.resize(P->ngrids);
    r_abs.resize(P->ngrids);
    b_abs.resize(P->ngrids);
    lin_r0_.assign(P->ngrids, 0.0);
    for (int g = 0; g < P->ngrids; ++g) {
      const double rr = host4[NV * g], bb = host4[NV * g + 1];
      r_abs[g] = std::sqrt(rr);
      b_abs[g] = std::sqrt(bb);
      r[g] = bb > 0.0 ? std::sqrt(rr) / std::sqrt(bb) : std::sqrt(rr);
      if (fuse) lin_r0_[g] = std::sqrt(host4[NV * g + 2]);
    }
    r0_ready_ = fuse;
  }

  void time_begin(int what = IMAU_TIME_SMOOTHER) {
    if (c->ktime != what) return;
    if (c->ev_used + 2 > c->ev.size()) {
      for (int q = 0; q < 256; ++q) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        c->ev.push_back(e);
      }
    }
    HIPCHK(hipEventRecord(c->ev[c->ev_used], c->s));
  }
  void time_end(int what = IMAU_TIME_SMOOTHER) {
    if (c->ktime != what) return;
    const std::vector<int> &act = (what == IMAU_TIME_VISC) ? act_pic : act_lin;
    for (int g = 0; g < P->ngrids; ++g)
      if (act[g]) c->sweep_cells += (long long)P->mem[g].nx * P->mem[g].ny;
    HIPCHK(hipEventRecord(c->ev[c->ev_used + 1], c->s));
    c->ev_used += 2;
  }
  void time_collect() {
    if (!c->ktime || !c->ev_used) return;
    HIPCHK(hipStreamSynchronize(c->s));
    for (size_t q = 0; q + 1 < c->ev_used; q += 2) {
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, c->ev[q], c->ev[q + 1]));
      c->sweep_ms += ms;
      c->sweep_launches += 1;
    }
    c->ev_used = 0;
  }

  // linear solve of every Picard-active member (src/utilities_module.f90:600-718)
  void linear_solve() {
    act_lin = act_pic;
    push_active(P->act_lin.p, act_lin);
    if (!r0_ready_) imau::init_x(L, P->F, P->F.xu0, P->F.xv0);
    if (cfg->DIVA_choice_matrix_solver == IMAU_SOLVER_SOR) sor_solve();
    else if (cfg->ksp_type == 1) gcr_solve();
    else krylov_solve();
    r0_ready_ = false;
    time_collect();
  }

  // reference SOR semantics (stop when the largest Gauss-Seidel update is
  // <= DIVA_SOR_tol, at most DIVA_SOR_nit sweeps, divergence guard), in a
  // multicolour ordering: colours u-red, u-black, v-red, v-black, boundary rows
  void sor_solve() {
    const int K = cfg->check_interval > 0 ? cfg->check_interval : 8;
    double om = sor_omega;
    int it = 0;
    std::vector<unsigned long long> rm(P->ngrids);
    while (it < sor_nit) {
      const bool checking = ((it + 1) % K == 0) || (it + 1 == sor_nit);
      if (checking) imau::fill_u64(c->s, P->resmax.p, P->ngrids, 0ull);
      time_begin();
      for (int col = 0; col < 5; ++col)
        imau::sor_color(L, kp, coef(), P->F.xu0, P->F.xv0, col, om, P->act_lin.p, checking, P->resmax.p);
      time_end();
      kcheck();
      ++it;
      for (int g = 0; g < P->ngrids; ++g)
        if (act_lin[g]) P->sweeps[g] += 1;
      if (!checking) continue;
      HIPCHK(hipMemcpyAsync(rm.data(), P->resmax.p, P->ngrids * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->s));
      HIPCHK(hipStreamSynchronize(c->s));
      bool any = false, diverged = false;
      for (int g = 0; g < P->ngrids; ++g) {
        if (!act_lin[g]) continue;
        double r;
        std::memcpy(&r, &rm[g], sizeof r);
        if (!std::isfinite(r) || (it > 100 && r > 1e3)) diverged = true;
        if (r <= sor_tol) act_lin[g] = 0;
        else any = true;
      }
      if (diverged) {
        // src/utilities_module.f90:697-712: lower omega, restart from zero
        om -= 0.1;
        if (om <= 0.1) fail(IMAU_ERR_DIVERGED, "SOR diverged even with omega <= 0.1");
        it = 0;
        act_lin = act_pic;
        HIPCHK(hipMemsetAsync(P->F.xu0, 0, (size_t)P->na * sizeof(double), c->s));
        HIPCHK(hipMemsetAsync(P->F.xv0, 0, (size_t)P->na * sizeof(double), c->s));
        any = true;
      }
      if (!any) break;
      push_active(P->act_lin.p, act_lin);
    }
  }

  // KSP semantics (||A x - b|| <= max(rtol * ref, abstol), ref = ||b|| under
  // the Picard stop rule, the initial residual under the residual stop rule):
  // BiCGStab, right-preconditioned by tile-wise symmetric Gauss-Seidel
  void krylov_solve() {
    const int G = P->ngrids;
    double *X[2] = {P->F.xu0, P->F.xv0};
    double *R[2] = {f2p(P, F_ru), f2p(P, F_rv)}, *RH[2] = {f2p(P, F_rhu), f2p(P, F_rhv)};
    double *Pp[2] = {f2p(P, F_pu), f2p(P, F_pv)}, *Vv[2] = {f2p(P, F_vu), f2p(P, F_vv)};
    double *PH[2] = {f2p(P, F_phu), f2p(P, F_phv)}, *SH[2] = {f2p(P, F_shu), f2p(P, F_shv)};
    double *T[2] = {f2p(P, F_tu), f2p(P, F_tv)};
    const imau::Coef cf = coef();
    const bool use_mg = cfg->mg_levels != 1 && !P->mg.empty();
    std::vector<double> out(2 * (size_t)G), target(G, 0.0);
    auto fetch = [&]() {
      HIPCHK(hipMemcpyAsync(out.data(), P->msum.p, 2 * G * sizeof(double), hipMemcpyDeviceToHost, c->s));
      HIPCHK(hipStreamSynchronize(c->s));
    };
    imau::bicg_init(L, kp, cf, X[0], X[1], R[0], R[1], RH[0], RH[1], Pp[0], Pp[1], Vv[0], Vv[1], P->act_lin.p,
                    P->partial.p);
    imau::bicg_scalar(c->s, G, P->tstart_d.p, P->partial.p, P->sc.p, P->msum.p, P->act_lin.p, 0);
    kcheck();
    fetch();
    bool any = false;
    for (int g = 0; g < G; ++g) {
      if (!act_lin[g]) continue;
      const double r0 = std::sqrt(out[2 * g]), bn = std::sqrt(out[2 * g + 1]);
      if (!std::isfinite(r0)) fail(IMAU_ERR_NAN, "member %d: non-finite initial linear residual", g);
      const double ref = (cfg->stop_rule == IMAU_STOP_RESIDUAL) ? r0 : bn;
      target[g] = std::max(petsc_rtol * ref, petsc_abstol);
      if (r0 <= target[g]) act_lin[g] = 0;
      else any = true;
    }
    if (!any) return;
    push_active(P->act_lin.p, act_lin);
    const bool fp32 = cfg->precision == IMAU_PREC_MIXED || cfg->precision == IMAU_PREC_MIXED_VISC;
    P->plan.fused = cfg->mg_smoother == 0;
    P->plan.periodic = false;
    for (int q = 0; q < 8; ++q) P->plan.periodic = P->plan.periodic || (&cfg->DIVA_boundary_BC_u_west)[q] == IMAU_BC_PERIODIC;
    if (use_mg) {
      const int every = cfg->mg_rebuild > 0 ? cfg->mg_rebuild : 1;
      imau::mg_setup(L, kp, cf, P->plan, fp32, mg_solves_ % every == 0, P->act_lin.p);
      ++mg_solves_;
    }
    auto precondition = [&](double *const *y, double *const *z) {
      if (use_mg) imau::mg_vcycle(L, kp, cf, P->plan, fp32, y[0], y[1], z[0], z[1], P->act_lin.p, &Solver::mg_mark, this);
      else {
        time_begin();
        imau::prec_tile(L, kp, cf, y[0], y[1], z[0], z[1], P->act_lin.p, 1);
        time_end();
      }
    };
    // Each member keeps its best iterate: BiCGStab's residual is not monotone,
    // and on nearly singular systems (the ice-free rows of choice_ice_margin
    // = 'BC') it can run away.  A member whose residual grows beyond
    // kRunaway times its best restarts from that iterate (r recomputed, the
    // BiCGStab scalars reset), at most kRestarts times; a member that ends
    // unconverged returns its best iterate, never one worse than its start.
    constexpr double kRunaway = 1e4;
    constexpr int kRestarts = 10;
    double *XB[2] = {f2p(P, F_xbu), f2p(P, F_xbv)};
    std::vector<double> best(G, INFINITY);
    std::vector<int> nrestart(G, 0);
    auto push_lin = [&](const std::vector<int> &which) {
      HIPCHK(hipStreamSynchronize(c->s));  // lin_flag_ is reused below
      lin_flag_ = which;
      HIPCHK(hipMemcpyAsync(P->lin_d.p, lin_flag_.data(), G * sizeof(int), hipMemcpyHostToDevice, c->s));
    };
    for (int g = 0; g < G; ++g)
      if (act_lin[g]) best[g] = std::sqrt(out[2 * g]);
    push_lin(act_lin);
    imau::copy2_members(L, X[0], X[1], XB[0], XB[1], P->lin_d.p);
    const int maxit = sor_nit > 0 ? sor_nit : 10000;
    for (int it = 0; it < maxit; ++it) {
      imau::bicg_p(L, R[0], R[1], Pp[0], Pp[1], Vv[0], Vv[1], P->sc.p, P->act_lin.p);
      precondition(Pp, PH);
      imau::apply(L, kp, cf, PH[0], PH[1], Vv[0], Vv[1], RH[0], RH[1], P->act_lin.p, P->partial.p);
      imau::bicg_scalar(c->s, G, P->tstart_d.p, P->partial.p, P->sc.p, P->msum.p, P->act_lin.p, 1);
      imau::bicg_s(L, R[0], R[1], Vv[0], Vv[1], P->sc.p, P->act_lin.p);
      precondition(R, SH);
      imau::apply(L, kp, cf, SH[0], SH[1], T[0], T[1], R[0], R[1], P->act_lin.p, P->partial.p);
      imau::bicg_scalar(c->s, G, P->tstart_d.p, P->partial.p, P->sc.p, P->msum.p, P->act_lin.p, 2);
      imau::bicg_xr(L, X[0], X[1], R[0], R[1], PH[0], PH[1], SH[0], SH[1], T[0], T[1], RH[0], RH[1], P->sc.p,
                    P->act_lin.p, P->partial.p);
      imau::bicg_scalar(c->s, G, P->tstart_d.p, P->partial.p, P->sc.p, P->msum.p, P->act_lin.p, 3);
      kcheck();
      fetch();
      any = false;
      bool improved = false, runaway = false;
      std::vector<int> imp(G, 0), rst(G, 0);
      for (int g = 0; g < G; ++g) {
        if (!act_lin[g]) continue;
        P->sweeps[g] += 2;
        const double rn = std::sqrt(out[2 * g + 1]);
        if (!std::isfinite(rn) && nrestart[g] >= kRestarts)
          fail(IMAU_ERR_NAN, "member %d: non-finite linear residual", g);
        if (rn <= target[g]) {
          act_lin[g] = 0;
          continue;
        }
        any = true;
        if (rn < best[g]) {
          best[g] = rn;
          imp[g] = 1;
          improved = true;
        } else if (!(rn <= kRunaway * best[g]) && nrestart[g] < kRestarts) {
          rst[g] = 1;
          ++nrestart[g];
          runaway = true;
        }
      }
      if (improved) {
        push_lin(imp);
        imau::copy2_members(L, X[0], X[1], XB[0], XB[1], P->lin_d.p);
      }
      if (runaway) {
        push_lin(rst);
        imau::copy2_members(L, XB[0], XB[1], X[0], X[1], P->lin_d.p);
        imau::bicg_init(L, kp, cf, X[0], X[1], R[0], R[1], RH[0], RH[1], Pp[0], Pp[1], Vv[0], Vv[1], P->lin_d.p,
                        P->partial.p);
        imau::bicg_scalar(c->s, G, P->tstart_d.p, P->partial.p, P->sc.p, P->msum.p, P->lin_d.p, 0);
      }
      kcheck();
      if (!any) break;
      push_active(P->act_lin.p, act_lin);
    }
    // members that end unconverged return their best iterate
    bool left = false;
    for (int g = 0; g < G; ++g) left = left || act_lin[g];
    if (left) {
      push_lin(act_lin);
      imau::copy2_members(L, XB[0], XB[1], X[0], X[1], P->lin_d.p);
      kcheck();
    }
  }

  // preconditioner application z = M^-1 y shared by the Krylov methods
  // V-cycles captured as HIP graphs, one per (operands, precision, smoother
  // form) seen in this solve: the kernel parameters, coefficients and levels
  // are fixed for the solve, the activity flags are read from device memory
  // at replay.  Released with the solver.
  struct VGraph {
    const double *yu, *yv;
    double *zu, *zv;
    bool fp32, fused, periodic;
    hipGraphExec_t exec;
  };
  std::vector<VGraph> vgraphs_;
  ~Solver() {
    // copies from the host flag vectors below and graph launches may still be
    // queued when a solve ends on an error
    (void)hipStreamSynchronize(c->s);
    for (VGraph &g : vgraphs_) (void)hipGraphExecDestroy(g.exec);
  }
  void precondition(const imau::Coef &cf, bool use_mg, bool fp32, const double *yu, const double *yv, double *zu,
                    double *zv) {
    if (use_mg && cfg->mg_graph && c->ktime != IMAU_TIME_SMOOTHER) {  // (the smoother timing records events)
      for (VGraph &g : vgraphs_)
        if (g.yu == yu && g.yv == yv && g.zu == zu && g.zv == zv && g.fp32 == fp32 && g.fused == P->plan.fused &&
            g.periodic == P->plan.periodic) {
          HIPCHK(hipGraphLaunch(g.exec, c->s));
          return;
        }
      hipGraph_t graph = nullptr;
      HIPCHK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
      imau::mg_vcycle(L, kp, cf, P->plan, fp32, yu, yv, zu, zv, P->act_lin.p, nullptr, nullptr);
      const hipError_t ec = hipStreamEndCapture(c->s, &graph);
      HIPCHK(ec);
      VGraph g{yu, yv, zu, zv, fp32, P->plan.fused, P->plan.periodic, nullptr};
      const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      HIPCHK(ei);
      vgraphs_.push_back(g);
      HIPCHK(hipGraphLaunch(g.exec, c->s));
      return;
    }
    if (use_mg) {
      imau::mg_vcycle(L, kp, cf, P->plan, fp32, yu, yv, zu, zv, P->act_lin.p, &Solver::mg_mark, this);
    } else {
      time_begin();
      imau::prec_tile(L, kp, cf, yu, yv, zu, zv, P->act_lin.p, 1);
      time_end();
    }
  }

  // KSP semantics as in krylov_solve, with right-preconditioned GCR restarted
  // every IMAU_GCR_M directions: per iteration z = M^-1 r, q = A z, q and z
  // orthogonalised against the stored directions, x += alpha z, r -= alpha q
  void gcr_solve() {
    const int G = P->ngrids;
    const long long na = P->na;
    if (!P->gcr_dirs.p) P->gcr_dirs.alloc((size_t)4 * IMAU_GCR_M * na);
    GList Q{}, Z{};
    double *Qd[2][IMAU_GCR_M], *Zd[2][IMAU_GCR_M];
    for (int k = 0; k < IMAU_GCR_M; ++k) {
      for (int c2 = 0; c2 < 2; ++c2) {
        Qd[c2][k] = P->gcr_dirs.p + (size_t)(2 * k + c2) * na;
        Zd[c2][k] = P->gcr_dirs.p + (size_t)(2 * IMAU_GCR_M + 2 * k + c2) * na;
      }
      Q.u[k] = Qd[0][k]; Q.v[k] = Qd[1][k];
      Z.u[k] = Zd[0][k]; Z.v[k] = Zd[1][k];
    }
    double *X[2] = {P->F.xu0, P->F.xv0};
    double *R[2] = {f2p(P, F_ru), f2p(P, F_rv)};
    const imau::Coef cf = coef();
    const bool use_mg = cfg->mg_levels != 1 && !P->mg.empty();
    const bool fp32 = cfg->precision == IMAU_PREC_MIXED || cfg->precision == IMAU_PREC_MIXED_VISC;
    std::vector<double> out(2 * (size_t)G), target(G, 0.0);
    auto fetch = [&]() {
      HIPCHK(hipMemcpyAsync(out.data(), P->msum.p, 2 * G * sizeof(double), hipMemcpyDeviceToHost, c->s));
      HIPCHK(hipStreamSynchronize(c->s));
    };
    // r = b - A x and the norms of r and b (the BiCGStab start kernel, r only),
    // unless the residual check of this Picard iteration prepared them
    if (!r0_ready_) {
      imau::bicg_init(L, kp, cf, X[0], X[1], R[0], R[1], nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                      P->act_lin.p, P->partial.p);
      imau::member_sum2(c->s, G, P->tstart_d.p, P->partial.p, P->msum.p);
      kcheck();
      fetch();
    } else {
      for (int g = 0; g < G; ++g) {
        out[2 * g] = lin_r0_[g] * lin_r0_[g];
        out[2 * g + 1] = b_abs[g] * b_abs[g];
      }
    }
    bool any = false;
    for (int g = 0; g < G; ++g) {
      if (!act_lin[g]) continue;
      const double r0 = std::sqrt(out[2 * g]), bn = std::sqrt(out[2 * g + 1]);
      if (!std::isfinite(r0)) fail(IMAU_ERR_NAN, "member %d: non-finite initial linear residual", g);
      const double ref = (cfg->stop_rule == IMAU_STOP_RESIDUAL) ? r0 : bn;
      target[g] = std::max(petsc_rtol * ref, petsc_abstol);
      if (r0 <= target[g]) act_lin[g] = 0;
      else any = true;
    }
    if (!any) return;
    push_active(P->act_lin.p, act_lin);
    P->plan.fused = cfg->mg_smoother == 0;
    P->plan.periodic = false;
    for (int q = 0; q < 8; ++q) P->plan.periodic = P->plan.periodic || (&cfg->DIVA_boundary_BC_u_west)[q] == IMAU_BC_PERIODIC;
    if (use_mg) {
      const int every = cfg->mg_rebuild > 0 ? cfg->mg_rebuild : 1;
      imau::mg_setup(L, kp, cf, P->plan, fp32, mg_solves_ % every == 0, P->act_lin.p);
      ++mg_solves_;
    }
    const int maxit = sor_nit > 0 ? sor_nit : 10000;
    int nj = 0;
    for (int it = 0; it < maxit; ++it) {
      precondition(cf, use_mg, fp32, R[0], R[1], Zd[0][nj], Zd[1][nj]);
      imau::apply(L, kp, cf, Zd[0][nj], Zd[1][nj], Qd[0][nj], Qd[1][nj], nullptr, nullptr, P->act_lin.p, nullptr);
      imau::gcr_dots(L, Qd[0][nj], Qd[1][nj], R[0], R[1], Q, nj, P->act_lin.p, P->partial.p);
      imau::gcr_scalar(c->s, G, P->tstart_d.p, P->partial.p, P->sc.p, P->act_lin.p, nj);
      imau::gcr_update(L, Q, Z, nj, X[0], X[1], R[0], R[1], P->sc.p, P->act_lin.p, P->partial.p);
      imau::member_sum2(c->s, G, P->tstart_d.p, P->partial.p, P->msum.p);
      kcheck();
      fetch();
      any = false;
      for (int g = 0; g < G; ++g) {
        if (!act_lin[g]) continue;
        P->sweeps[g] += 1;
        const double rn = std::sqrt(out[2 * g]);
        if (!std::isfinite(rn)) fail(IMAU_ERR_NAN, "member %d: non-finite linear residual", g);
        if (rn <= target[g]) act_lin[g] = 0;
        else any = true;
      }
      if (!any) break;
      push_active(P->act_lin.p, act_lin);
      nj = (nj + 1 == IMAU_GCR_M) ? 0 : nj + 1;
    }
  }

  // ----- Anderson acceleration of the Picard map (host side) -----
  long long and_count_ = 0;              // Anderson steps taken in this solve
  std::vector<long long> and_since_;     // per member: step of its last restart
  std::vector<double> and_gram_;         // per member: Gram matrix (F_s, F_t) of the slots
  std::vector<double> and_w_h_;          // per member: a_s, b_s

  void and_init(AMix &mix) {
    const long long na = P->na;
    if (!P->and_slots.p) P->and_slots.alloc((size_t)4 * IMAU_AND_SLOTS * na);
    if (!P->and_w.p) P->and_w.alloc((size_t)2 * IMAU_AND_SLOTS * P->ngrids);
    for (int q = 0; q < IMAU_AND_SLOTS; ++q) {
      mix.U.u[q] = P->and_slots.p + (size_t)(4 * q + 0) * na;
      mix.U.v[q] = P->and_slots.p + (size_t)(4 * q + 1) * na;
      mix.F.u[q] = P->and_slots.p + (size_t)(4 * q + 2) * na;
      mix.F.v[q] = P->and_slots.p + (size_t)(4 * q + 3) * na;
    }
    mix.ns = 0;
    mix.w = nullptr;
  }

  // Store slot k = (U_k, F_k), update the Gram matrices, solve the per-member
  // least-squares problem min |F_k - sum_i gamma_i dF_i| over the member's
  // valid slots and set u_new = G_k - sum_i gamma_i dG_i (G_s = U_s + F_s).
  void and_step(AMix &mix, int m) {
    const int G = P->ngrids, S = m + 1;
    const int slot = (int)(and_count_ % S);
    mix.ns = (int)std::min<long long>(and_count_ + 1, S);
    mix.w = nullptr;
    imau::and_prep(L, kp, P->F, P->F.xu0, P->F.xv0, mix, slot, P->act_pic.p, P->partial.p);
    imau::member_sum_and(c->s, G, P->tstart_d.p, P->partial.p, P->msum.p);
    kcheck();
    std::vector<double> dots((size_t)IMAU_AND_SLOTS * G);
    HIPCHK(hipMemcpyAsync(dots.data(), P->msum.p, dots.size() * sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    and_w_h_.assign((size_t)2 * IMAU_AND_SLOTS * G, 0.0);
    const double rel = cfg->DIVA_visc_it_relax;
    for (int g = 0; g < G; ++g) {
      double *W = &and_w_h_[(size_t)2 * IMAU_AND_SLOTS * g];
      double *Gm = &and_gram_[(size_t)IMAU_AND_SLOTS * IMAU_AND_SLOTS * g];
      for (int q = 0; q < mix.ns; ++q) {
        Gm[slot * IMAU_AND_SLOTS + q] = dots[(size_t)IMAU_AND_SLOTS * g + q];
        Gm[q * IMAU_AND_SLOTS + slot] = dots[(size_t)IMAU_AND_SLOTS * g + q];
      }
      // the member's slots, oldest first
      const int nv = (int)std::min<long long>(and_count_ - and_since_[g] + 1, S);
      int sl[IMAU_AND_SLOTS];
      for (int i = 0; i < nv; ++i) sl[i] = (int)((and_count_ - (nv - 1) + i) % S);
      bool mixed = false;
      if (nv >= 2) {
        const int n = nv - 1;
        double M[IMAU_AND_SLOTS][IMAU_AND_SLOTS + 1];
        auto gr = [&](int a, int b) { return Gm[sl[a] * IMAU_AND_SLOTS + sl[b]]; };
        double tr = 0.0;
        for (int a = 0; a < n; ++a) {
          for (int b = 0; b < n; ++b) M[a][b] = gr(a + 1, b + 1) - gr(a + 1, b) - gr(a, b + 1) + gr(a, b);
          M[a][n] = gr(a + 1, n) - gr(a, n);
          tr += M[a][a];
        }
        for (int a = 0; a < n; ++a) M[a][a] += 1e-10 * tr;
        bool ok = tr > 0.0 && std::isfinite(tr);
        for (int a = 0; a < n && ok; ++a) {  // Gaussian elimination with partial pivoting
          int piv = a;
          for (int b = a + 1; b < n; ++b)
            if (std::fabs(M[b][a]) > std::fabs(M[piv][a])) piv = b;
          if (!(std::fabs(M[piv][a]) > 1e-14 * tr)) { ok = false; break; }
          for (int cidx = 0; cidx <= n; ++cidx) std::swap(M[a][cidx], M[piv][cidx]);
          for (int b = a + 1; b < n; ++b) {
            const double f = M[b][a] / M[a][a];
            for (int cidx = a; cidx <= n; ++cidx) M[b][cidx] -= f * M[a][cidx];
          }
        }
        double gam[IMAU_AND_SLOTS] = {0, 0, 0, 0};
        if (ok) {
          for (int a = n - 1; a >= 0; --a) {
            double v = M[a][n];
            for (int b = a + 1; b < n; ++b) v -= M[a][b] * gam[b];
            gam[a] = v / M[a][a];
          }
          for (int a = 0; a < n; ++a) ok = ok && std::isfinite(gam[a]);
        }
        if (ok) {
          double w[IMAU_AND_SLOTS] = {0, 0, 0, 0};
          w[n] = 1.0;
          for (int a = 0; a < n; ++a) {
            w[a + 1] -= gam[a];
            w[a] += gam[a];
          }
          for (int i = 0; i < nv; ++i) {
            W[sl[i]] = w[i];
            W[IMAU_AND_SLOTS + sl[i]] = cfg->anderson_beta * w[i];
          }
          mixed = true;
        }
      }
      if (!mixed) {  // relaxation step u_k + rel (G(u_k) - u_k)
        W[slot] = 1.0;
        W[IMAU_AND_SLOTS + slot] = relaxed_ ? relax_h_[g] : rel;
        and_since_[g] = and_count_;
      }
    }
    HIPCHK(hipMemcpyAsync(P->and_w.p, and_w_h_.data(), and_w_h_.size() * sizeof(double), hipMemcpyHostToDevice, c->s));
    mix.w = P->and_w.p;
    ++and_count_;
  }

  static void mg_mark(void *self, int which) {
    Solver *S = static_cast<Solver *>(self);
    if (which == 0) S->time_begin();
    else S->time_end();
  }

  int sor_nit = 10000;
  double sor_tol = 2.5, sor_omega = 1.3, petsc_rtol = 0.01, petsc_abstol = 2.5;
  std::vector<double> relax_h_;
  long long mg_solves_ = 0;  // multigrid-preconditioned linear solves in this solve

  // stall restarts (imau_cfg.stall_window): called with the residuals r of the
  // current iterate, before the convergence check of this Picard iteration
  std::vector<std::vector<double>> stall_hist_;
  std::vector<double> *prev_r_ = nullptr;
  std::vector<int> restart_flag_, stop_flag_, lin_flag_;
  bool relaxed_ = false;  // per-member relaxation factors in use (adaptive mode or a restart)
  template <class VB>
  void stall_restarts(std::vector<double> &r, bool diva, VB &visc_beta, float *cur32) {
    const int G = P->ngrids, Wn = cfg->stall_window;
    std::vector<int> rs(G, 0);
    bool any = false;
    for (int g = 0; g < G; ++g) {
      if (!act_pic[g] || !(r[g] > cfg->nonlinear_rtol)) continue;
      std::vector<double> &h = stall_hist_[g];
      h.push_back(r[g]);
      if (P->restarts[g] > 0) continue;  // a member restarts at most once
      const int n = (int)h.size();
      // the first residual of zero velocities is 1: a start no better than that is dropped at once
      const bool stall = (n == 1 && !(r[g] < 1.0)) || (n > Wn && r[g] > h[n - 1 - Wn] / cfg->stall_drop);
      if (stall) { rs[g] = 1; any = true; }
    }
    if (!any) return;
    HIPCHK(hipStreamSynchronize(c->s));  // restart_flag_ is reused below
    restart_flag_ = rs;
    HIPCHK(hipMemcpyAsync(P->restart_d.p, restart_flag_.data(), G * sizeof(int), hipMemcpyHostToDevice, c->s));
    imau::reset_members(L, kp, P->F, imau::ResetSrc{P->u0, P->v0, P->tbx0, P->tby0, P->visc0}, P->visc_cur, cur32,
                        P->restart_d.p);
    // coefficients and residual of the restarted members at their new state; the
    // other members keep theirs (the second evaluation leaves them inactive)
    const std::vector<double> r_keep = r, ra_keep = r_abs, rb_keep = b_abs, l0_keep = lin_r0_;
    const bool ready = r0_ready_;
    push_active(P->act_pic.p, rs);
    visc_beta(false);
    imau::coef(L, kp, P->F, diva, P->act_pic.p);
    kcheck();
    std::vector<double> r2;
    residual(r2);
    for (int g = 0; g < G; ++g) {
      if (rs[g]) {
        r[g] = r2[g];
        P->restarts[g] = 1;
        stall_hist_[g].assign(1, r2[g]);
        relax_h_[g] = cfg->stall_relax > 0.0 ? cfg->stall_relax : cfg->DIVA_visc_it_relax;
        and_since_[g] = and_count_;
        (*prev_r_)[g] = -1.0;
      } else {
        r[g] = r_keep[g];
        r_abs[g] = ra_keep[g];
        b_abs[g] = rb_keep[g];
        if (!l0_keep.empty()) lin_r0_[g] = l0_keep[g];
      }
    }
    r0_ready_ = ready;
    relaxed_ = true;
    push_active(P->act_pic.p, act_pic);
  }

  void picard(bool diva) {
    const int G = P->ngrids;
    act_pic.assign(G, 0);
    for (int g = 0; g < G; ++g) act_pic[g] = P->sheet_any[g];
    if (!diva && cfg->choice_sliding_law == IMAU_SLID_NO_SLIDING) std::fill(act_pic.begin(), act_pic.end(), 0);
    reserve_buffers(P, cfg);
    std::fill(P->picard_its.begin(), P->picard_its.end(), 0);
    std::fill(P->restarts.begin(), P->restarts.end(), 0);
    std::fill(P->sweeps.begin(), P->sweeps.end(), 0);
    std::fill(P->nl_resid.begin(), P->nl_resid.end(), NAN);
    imau::setup1(L, kp, P->F, P->yg.p);
    imau::setup2(L, kp, P->F);
    imau::mg_plain_flags(L, P->F.rt_u, P->F.rt_v, P->plan);
    if (P->err.p) {  // :176-177 / :303-304, after the early return of a member without grounded ice
      push_active(P->act_pic.p, act_pic);
      imau::fill2_members(L, P->err.p, P->err.p + P->na, 1e5, P->act_pic.p);
      for (int g = 0; g < G; ++g)
        if (act_pic[g]) P->err_slot[g] = 0;
    }
    kcheck();
    if (!P->plan.levels.empty()) {  // the length of the non-plain list, once per solve
      HIPCHK(hipMemcpyAsync(&P->plan.nlist, P->plan.nlist_d, sizeof(int), hipMemcpyDeviceToHost, c->s));
      HIPCHK(hipStreamSynchronize(c->s));
    }
    std::vector<double> r;
    // adaptive relaxation: per-member factor, raised while the FP64 residual falls
    const bool adaptive = cfg->picard_relax_max > 0.0 && cfg->stop_rule == IMAU_STOP_RESIDUAL;
    relax_h_.assign(G, cfg->DIVA_visc_it_relax);
    std::vector<double> prev_r(G, -1.0);
    prev_r_ = &prev_r;
    // stall restarts: the FP64 residual history of every member since its start
    const bool stall_on = cfg->stop_rule == IMAU_STOP_RESIDUAL && cfg->stall_window > 0;
    stall_hist_.assign(G, std::vector<double>());
    relaxed_ = adaptive;
    // Anderson acceleration state
    const int and_m = (cfg->stop_rule == IMAU_STOP_RESIDUAL) ? cfg->anderson_m : 0;
    AMix mix{};
    if (and_m > 0) and_init(mix);
    and_count_ = 0;
    and_since_.assign(G, 0);
    and_gram_.assign((size_t)G * IMAU_AND_SLOTS * IMAU_AND_SLOTS, 0.0);
    // mixed precision: the 3D viscosity and the rate-factor term in FP32
    const bool v32 = cfg->precision == IMAU_PREC_MIXED_VISC;
    const long long n3 = P->na * (long long)P->nz;
    float *cur32 = nullptr, *oth32 = nullptr;
    if (v32) {
      if (!P->visc32.p) P->visc32.alloc((size_t)3 * n3);
      cur32 = P->visc32.p;
      oth32 = P->visc32.p + n3;
      imau::to_f32(c->s, P->F.B3, P->visc32.p + 2 * n3, n3);
      imau::to_f32(c->s, P->visc_cur, cur32, n3);
    }
    // viscosity, F2 and beta_eff at the current velocities (A4-A8)
    auto visc_beta = [&](bool timed) {
      if (timed) time_begin(IMAU_TIME_VISC);
      if (v32) imau::visc3d32(L, LV(), kp, P->F, diva, cur32, oth32, P->visc32.p + 2 * n3, P->act_pic.p);
      else imau::visc3d(L, LV(), kp, P->F, diva, P->visc_cur, P->visc_oth, P->act_pic.p);
      if (timed) time_end(IMAU_TIME_VISC);
      imau::beta(L, kp, P->F, diva, P->xg.p, P->yg.p, P->act_pic.p);
      if (v32) std::swap(cur32, oth32);
      else std::swap(P->visc_cur, P->visc_oth);
    };
    // warm start given as velocities only: bring viscosity and basal stress to them
    if (diva && (cfg->warm_visc_iters > 0 || cfg->warm_column_iters > 0)) {
      push_active(P->act_pic.p, act_pic);
      if (cfg->warm_column_iters > 0) {
        if (v32)
          imau::visc_column32(L, LV(), kp, P->F, cur32, P->visc32.p + 2 * n3, P->xg.p, P->yg.p,
                              cfg->warm_column_iters, P->act_pic.p);
        else
          imau::visc_column(L, LV(), kp, P->F, P->visc_cur, P->xg.p, P->yg.p, cfg->warm_column_iters,
                            P->act_pic.p);
        imau::coef(L, kp, P->F, diva, P->act_pic.p);  // beta_eff on cx / cy
        imau::taub(L, P->F, P->act_pic.p);
      }
      for (int w = 0; w < cfg->warm_visc_iters; ++w) {
        visc_beta(false);
        imau::coef(L, kp, P->F, diva, P->act_pic.p);  // beta_eff on cx / cy
        imau::taub(L, P->F, P->act_pic.p);
      }
      kcheck();
    }
    for (int it = 1;; ++it) {
      bool any = false;
      for (int g = 0; g < G; ++g) any = any || act_pic[g];
      if (!any) break;
      push_active(P->act_pic.p, act_pic);
      visc_beta(true);
      imau::coef(L, kp, P->F, diva, P->act_pic.p);
      kcheck();
      if (cfg->stop_rule == IMAU_STOP_RESIDUAL) {
        residual(r);
        if (stall_on) stall_restarts(r, diva, visc_beta, v32 ? cur32 : nullptr);
        bool left = false, stopped = false;
        std::vector<int> stop(G, 0);
        for (int g = 0; g < G; ++g) {
          if (!act_pic[g]) continue;
          P->nl_resid[g] = r[g];
          if (r[g] <= cfg->nonlinear_rtol || it > cfg->DIVA_visc_it_nit) {
            act_pic[g] = 0;
            stop[g] = 1;
            stopped = true;
          } else {
            left = true;
          }
          if (adaptive) {
            if (prev_r[g] >= 0.0)
              relax_h_[g] = (r[g] < prev_r[g]) ? std::min(cfg->picard_relax_max, 1.25 * relax_h_[g])
                                               : cfg->DIVA_visc_it_relax;
            prev_r[g] = r[g];
          }
          if (and_m > 0) {  // restart the member's history after a 1.5x residual rise
            if (prev_r[g] >= 0.0 && r[g] > cfg->anderson_restart * prev_r[g]) and_since_[g] = and_count_;
            prev_r[g] = r[g];
          }
        }
        if (stopped) {
          // a member that stops returns the state whose residual was evaluated:
          // its velocities and basal stress with the viscosity they were
          // computed with (the check's viscosity is that one's next iterate)
          HIPCHK(hipStreamSynchronize(c->s));  // stop_flag_ is reused below
          stop_flag_ = stop;
          HIPCHK(hipMemcpyAsync(P->stop_d.p, stop_flag_.data(), G * sizeof(int), hipMemcpyHostToDevice, c->s));
          const imau::Launch L3 = L;
          if (v32) imau::copy3_members(L3, P->nz, oth32, cur32, P->stop_d.p);
          else imau::copy3_members(L3, P->nz, P->visc_oth, P->visc_cur, P->stop_d.p);
          kcheck();
        }
        if (!left) break;
        push_active(P->act_pic.p, act_pic);
      }
      linear_solve();
      if (relaxed_)  // relax_h_ is not modified before the next stream synchronisation
        HIPCHK(hipMemcpyAsync(P->relax_d.p, relax_h_.data(), G * sizeof(double), hipMemcpyHostToDevice, c->s));
      if (and_m > 0) and_step(mix, and_m);
      if (P->err.p) {  // this iteration's error estimate goes to slot it & 1
        P->F.errt_u = P->err.p + (size_t)(it & 1) * 2 * P->na;
        P->F.errt_v = P->F.errt_u + P->na;
      }
      imau::post(L, kp, P->F, P->F.xu0, P->F.xv0, P->act_pic.p, diva, P->partial.p, relaxed_ ? P->relax_d.p : nullptr,
                 mix);
      imau::member_sum3(c->s, G, P->tstart_d.p, P->partial.p, P->msum.p);
      kcheck();
      HIPCHK(hipMemcpyAsync(host4.data(), P->msum.p, 3 * G * sizeof(double), hipMemcpyDeviceToHost, c->s));
      HIPCHK(hipStreamSynchronize(c->s));
      const std::vector<int> ran = act_pic;
      for (int g = 0; g < G; ++g) {
        if (!act_pic[g]) continue;
        P->picard_its[g] = it;
        const double r1 = host4[3 * g], r2 = host4[3 * g + 1], n = host4[3 * g + 2];
        double res = 0.0;
        if (n > 0.0) res = 2.0 * std::sqrt(r1) / std::max(std::sqrt(r2), 1e-8);
        if (!std::isfinite(r1) || !std::isfinite(r2)) fail(IMAU_ERR_NAN, "member %d: non-finite velocity in Picard iteration %d", g, it);
        P->resid_uv[g] = res;
        if (cfg->stop_rule == IMAU_STOP_PICARD && (res < cfg->DIVA_visc_it_norm_dUV_tol || it >= cfg->DIVA_visc_it_nit))
          act_pic[g] = 0;
      }
      if (P->err.p)  // :355: the estimate of an iteration that does not end the loop is kept
        for (int g = 0; g < G; ++g)
          if (ran[g] && (cfg->stop_rule != IMAU_STOP_PICARD || act_pic[g])) P->err_slot[g] = it & 1;
    }
    if (v32) imau::to_f64(c->s, cur32, P->visc_cur, n3);  // the solve's viscosity output in FP64
  }

  void finish(int dyn) {
    time_collect();  // event pairs of the last Picard iteration (it ends before a linear solve)
    if (dyn == IMAU_DYN_DIVA) {
      imau::F1(L, kp, P->F, P->visc_cur, P->O);
      imau::uv3d(L, kp, P->F, P->O);
      imau::final_a(L, kp, P->F, P->O, dyn, P->F.u, P->F.v, true);
    } else {
      imau::compose(L, kp, P->F, P->O, dyn == IMAU_DYN_SIA ? 1 : (dyn == IMAU_DYN_SSA ? 2 : 3));
      imau::final_a(L, kp, P->F, P->O, dyn, P->O.uvav, P->O.vvav, dyn != IMAU_DYN_SSA);
    }
    kcheck();
  }
};

}  // namespace

namespace {
enum Kind3 { SOLVE_SIA = 1, SOLVE_SSA = 2, SOLVE_DIVA = 4 };

int solve_kind(imau_ctx *ctx, imau_problem *P, const imau_cfg *cfg, int kinds) {
  ABI_TRY
  if (!ctx || !P) fail(IMAU_ERR_ARG, "NULL argument");
  check_cfg(cfg);
  if (cfg->nz != P->nz) fail(IMAU_ERR_CONFIG, "cfg nz=%d differs from the problem's nz=%d", cfg->nz, P->nz);
  check_ismip_e(P, cfg);
  if (cfg->include_basal_freezing && !P->has_ti) fail(IMAU_ERR_CONFIG, "include_basal_freezing needs Ti_a/Ti_pmp_a at create");
  const int dyn = cfg->choice_ice_dynamics;
  auto t0 = std::chrono::steady_clock::now();
  Solver S(ctx, P, cfg);
  S.sor_nit = P->sor_nit_;
  S.sor_tol = P->sor_tol_;
  S.sor_omega = P->sor_omega_;
  S.petsc_rtol = P->petsc_rtol_;
  S.petsc_abstol = P->petsc_abstol_;
  if (kinds & SOLVE_SIA) {
    if (dyn != IMAU_DYN_SIA && dyn != IMAU_DYN_SIA_SSA)
      fail(IMAU_ERR_CONFIG, "solve_SIA: choice_ice_dynamics is not SIA or SIA/SSA");
    imau::sia(S.L, S.kp, P->F, P->Acx3, P->Acy3, P->O);
    S.kcheck();
    S.finish(dyn);
  }
  if (kinds & SOLVE_SSA) {
    if (dyn != IMAU_DYN_SSA && dyn != IMAU_DYN_SIA_SSA)
      fail(IMAU_ERR_CONFIG, "solve_SSA: choice_ice_dynamics is not SSA or SIA/SSA");
    S.picard(false);
    S.finish(dyn);
  }
  if (kinds & SOLVE_DIVA) {
    if (dyn != IMAU_DYN_DIVA) fail(IMAU_ERR_CONFIG, "solve_DIVA: choice_ice_dynamics is not DIVA");
    S.picard(true);
    S.finish(dyn);
  }
  HIPCHK(hipStreamSynchronize(ctx->s));
  P->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return IMAU_OK;
  ABI_CATCH(ctx)
}

// download a k-plane (k = 0 surface / nz-1 base) of a cx/cy 3D field as a 2D field
void down_plane(imau_ctx *c, imau_problem *P, int g, Kind k, double *dst, const double *src3, int kk) {
  if (!dst) return;
  const Member &m = P->mem[g];
  int ni = m.nx, nj = m.ny;
  if (k == KCX) ni = m.nx - 1;
  if (k == KCY) nj = m.ny - 1;
  const double *src = src3 + m.off3 + (long long)kk * m.nx * m.ny;
  HIPCHK(hipMemcpy2DAsync(dst, (size_t)nj * sizeof(double), src, (size_t)m.ny * sizeof(double), (size_t)nj * sizeof(double),
                          ni, hipMemcpyDeviceToHost, c->s));
}
}  // namespace

extern "C" {

int imau_vel_problem_solve(imau_ctx *ctx, imau_problem *P, const imau_cfg *cfg) {
  if (!cfg) return IMAU_ERR_ARG;
  const int dyn = cfg->choice_ice_dynamics;
  int kinds = dyn == IMAU_DYN_DIVA ? SOLVE_DIVA
            : dyn == IMAU_DYN_SSA ? SOLVE_SSA
            : dyn == IMAU_DYN_SIA ? SOLVE_SIA
            : dyn == IMAU_DYN_SIA_SSA ? (SOLVE_SIA | SOLVE_SSA) : 0;
  if (!kinds) {
    if (ctx) ctx->err = "unknown choice_ice_dynamics";
    return IMAU_ERR_CONFIG;
  }
  return solve_kind(ctx, P, cfg, kinds);
}

int imau_vel_problem_download(imau_ctx *ctx, imau_problem *P, imau_ice_io *ios) {
  ABI_TRY
  if (!ctx || !P || !ios) fail(IMAU_ERR_ARG, "NULL argument");
  const size_t d = sizeof(double), q = sizeof(int32_t);
  std::vector<Down2Req> batch;
  struct Collect {  // the 2D fields below are collected, then copied by flush_down2
    imau_ctx *c;
    ~Collect() { c->batch = nullptr; }
  } collect{ctx};
  ctx->batch = &batch;
  const int dyn = P->dyn_uploaded;
  const bool ssa = uses_ssa_field(dyn);
  for (int g = 0; g < P->ngrids; ++g) {
    imau_ice_io &io = ios[g];
    if (dyn == IMAU_DYN_DIVA) {
      down2(ctx, P, g, KCX, io.u_vav_cx, P->F.u, d);
      down2(ctx, P, g, KCY, io.v_vav_cy, P->F.v, d);
    } else {
      down2(ctx, P, g, KCX, io.u_vav_cx, P->O.uvav, d);
      down2(ctx, P, g, KCY, io.v_vav_cy, P->O.vvav, d);
    }
    if (ssa) {
      down2(ctx, P, g, KCX, io.u_SSA_cx, P->F.u, d);
      down2(ctx, P, g, KCY, io.v_SSA_cy, P->F.v, d);
    }
    if (dyn == IMAU_DYN_DIVA) {
      down2(ctx, P, g, KCX, io.taub_cx, P->F.taub_cx, d);
      down2(ctx, P, g, KCY, io.taub_cy, P->F.taub_cy, d);
    }
    if (dyn != IMAU_DYN_SIA) down3(ctx, P, g, KA, io.visc_eff_3D_a, P->visc_cur);
    down3(ctx, P, g, KCX, io.u_3D_cx, P->O.u3);
    down3(ctx, P, g, KCY, io.v_3D_cy, P->O.v3);
    if (dyn == IMAU_DYN_SIA || dyn == IMAU_DYN_SIA_SSA) {
      down3(ctx, P, g, KCX, io.u_3D_SIA_cx, P->O.u3sia);
      down3(ctx, P, g, KCY, io.v_3D_SIA_cy, P->O.v3sia);
    }
    down3(ctx, P, g, KA, io.u_3D_a, P->O.u3a);
    down3(ctx, P, g, KA, io.v_3D_a, P->O.v3a);
    down3(ctx, P, g, KA, io.w_3D_a, P->O.w3);
    if (dyn == IMAU_DYN_DIVA) down3(ctx, P, g, KA, io.F1_3D_a, P->O.F1);
    down2(ctx, P, g, KA, io.u_vav_a, P->O.u_vav_a, d);
    down2(ctx, P, g, KA, io.v_vav_a, P->O.v_vav_a, d);
    down2(ctx, P, g, KA, io.u_surf_a, P->O.u_surf_a, d);
    down2(ctx, P, g, KA, io.v_surf_a, P->O.v_surf_a, d);
    down2(ctx, P, g, KA, io.u_base_a, P->O.u_base_a, d);
    down2(ctx, P, g, KA, io.v_base_a, P->O.v_base_a, d);
    down2(ctx, P, g, KA, io.uabs_vav_a, P->O.uabs_vav, d);
    down2(ctx, P, g, KA, io.uabs_surf_a, P->O.uabs_surf, d);
    down2(ctx, P, g, KA, io.uabs_base_a, P->O.uabs_base, d);
    down2(ctx, P, g, KA, io.w_surf_a, P->O.w_surf, d);
    down2(ctx, P, g, KA, io.w_base_a, P->O.w_base, d);
    down2(ctx, P, g, KA, io.R_shear, P->O.R_shear, d);
    down_plane(ctx, P, g, KCX, io.u_surf_cx, P->O.u3, 0);
    down_plane(ctx, P, g, KCY, io.v_surf_cy, P->O.v3, 0);
    down2(ctx, P, g, KCX, io.u_base_cx, P->F.u_base, d);
    down2(ctx, P, g, KCY, io.v_base_cy, P->F.v_base, d);
    if (dyn != IMAU_DYN_SIA) {
      down2(ctx, P, g, KA, io.beta_a, P->F.beta_a, d);
      down2(ctx, P, g, KA, io.beta_eff_a, P->F.beta_eff_a, d);
      down2(ctx, P, g, KCX, io.beta_eff_cx, P->F.beta_eff_cx, d);
      down2(ctx, P, g, KCY, io.beta_eff_cy, P->F.beta_eff_cy, d);
      down2(ctx, P, g, KCX, io.taudx_cx, P->F.taudx, d);
      down2(ctx, P, g, KCY, io.taudy_cy, P->F.taudy, d);
      down2(ctx, P, g, KA, io.N_a, P->F.N_a, d);
      down2(ctx, P, g, KB, io.N_b, P->F.N_b, d);
      down2(ctx, P, g, KCX, io.N_cx, P->F.N_cx, d);
      down2(ctx, P, g, KCY, io.N_cy, P->F.N_cy, d);
      down2(ctx, P, g, KA, io.visc_eff_int_a, P->F.visc_int, d);
      down2(ctx, P, g, KA, io.F2_a, P->F.F2, d);
      down2(ctx, P, g, KA, io.f_grnd_a, P->F.fg_a, d);
      down2(ctx, P, g, KCX, io.f_grnd_cx, P->F.fg_cx, d);
      down2(ctx, P, g, KCY, io.f_grnd_cy, P->F.fg_cy, d);
      down2(ctx, P, g, KB, io.f_grnd_b, P->F.fg_b, d);
      down2(ctx, P, g, KA, io.overburden_pressure_a, P->F.overb, d);
      down2(ctx, P, g, KA, io.pore_water_pressure_a, P->F.pore, d);
      down2(ctx, P, g, KA, io.Neff_a, P->F.Neff, d);
      down2(ctx, P, g, KA, io.beta_sq_a, P->F.beta_sq, d);
      down2(ctx, P, g, KA, io.alpha_sq_a, P->F.alpha_sq, d);
      down2(ctx, P, g, KA, io.phi_fric_a, P->F.phi_fric, d);
      down2(ctx, P, g, KA, io.tauc_a, P->F.tauc, d);
      down2(ctx, P, g, KA, io.DIVA_mask_combi, P->F.combi, q);
      down2(ctx, P, g, KA, io.DIVA_isfront_inner, P->F.fin, q);
      down2(ctx, P, g, KA, io.DIVA_isfront_outer, P->F.fout, q);
      down2(ctx, P, g, KCX, io.u_cx_prev, P->F.u_prev, d);
      if (P->err.p) {
        const double *e = P->err.p + (size_t)P->err_slot[g] * 2 * P->na;
        down2(ctx, P, g, KCX, io.DIVA_err_cx, e, d);
        down2(ctx, P, g, KCY, io.DIVA_err_cy, e + P->na, d);
      }
      down2(ctx, P, g, KCY, io.v_cy_prev, P->F.v_prev, d);
    }
    io.stat_picard_iterations = P->picard_its[g];
    io.stat_linear_sweeps = P->sweeps[g];
    io.stat_resid_UV = P->resid_uv[g];
    io.stat_nonlinear_residual = P->nl_resid[g];
    io.stat_seconds = P->seconds;
  }
  ctx->batch = nullptr;
  flush_down2(ctx, P, batch);
  HIPCHK(hipStreamSynchronize(ctx->s));
  // the reference sets u, v = 0 and returns for grids without grounded ice (:276-285)
  for (int g = 0; g < P->ngrids; ++g) {
    if (P->sheet_any[g]) continue;
    const Member &m = P->mem[g];
    imau_ice_io &io = ios[g];
    double *u = (dyn == IMAU_DYN_DIVA) ? io.u_vav_cx : io.u_SSA_cx;
    double *v = (dyn == IMAU_DYN_DIVA) ? io.v_vav_cy : io.v_SSA_cy;
    if (u) std::memset(u, 0, (size_t)m.ny * (m.nx - 1) * sizeof(double));
    if (v) std::memset(v, 0, (size_t)(m.ny - 1) * m.nx * sizeof(double));
  }
  return IMAU_OK;
  ABI_CATCH(ctx)
}

static int solve_convenience(imau_ctx *ctx, int ngrids, const imau_grid *grids, const imau_cfg *cfg, imau_ice_io *ios,
                             int kinds) {
  if (!ctx) return IMAU_ERR_ARG;
  imau_problem *P = nullptr;
  int rc = imau_vel_problem_create(ctx, ngrids, grids, cfg, ios, &P);
  if (rc) return rc;
  rc = solve_kind(ctx, P, cfg, kinds);
  if (!rc) rc = imau_vel_problem_download(ctx, P, ios);
  imau_vel_problem_destroy(P);
  return rc;
}

int imau_vel_solve_diva(imau_ctx *ctx, const imau_grid *grid, const imau_cfg *cfg, imau_ice_io *io) {
  return solve_convenience(ctx, 1, grid, cfg, io, SOLVE_DIVA);
}
int imau_vel_solve_ssa(imau_ctx *ctx, const imau_grid *grid, const imau_cfg *cfg, imau_ice_io *io) {
  return solve_convenience(ctx, 1, grid, cfg, io, SOLVE_SSA);
}
int imau_vel_solve_sia(imau_ctx *ctx, const imau_grid *grid, const imau_cfg *cfg, imau_ice_io *io) {
  return solve_convenience(ctx, 1, grid, cfg, io, SOLVE_SIA);
}
int imau_vel_solve_diva_batch(imau_ctx *ctx, int32_t ngrids, const imau_grid *grids, const imau_cfg *cfg,
                              imau_ice_io *ios) {
  return solve_convenience(ctx, ngrids, grids, cfg, ios, SOLVE_DIVA);
}

int imau_vel_nonlinear_residual(imau_ctx *ctx, const imau_grid *grid, const imau_cfg *cfg, imau_ice_io *io,
                                double *residual) {
  imau_problem *P = nullptr;
  struct Release {
    imau_problem *&p;
    ~Release() { imau_vel_problem_destroy(p); }
  } release{P};
  ABI_TRY
  if (!ctx || !grid || !cfg || !io || !residual) fail(IMAU_ERR_ARG, "NULL argument");
  int rc = imau_vel_problem_create(ctx, 1, grid, cfg, io, &P);
  if (rc) return rc;
  const bool diva = cfg->choice_ice_dynamics == IMAU_DYN_DIVA;
  Solver S(ctx, P, cfg);
  imau::setup1(S.L, S.kp, P->F, P->yg.p);
  imau::setup2(S.L, S.kp, P->F);
  std::vector<int> all(1, 1);
  S.push_active(P->act_pic.p, all);
  imau::visc3d(S.L, S.LV(), S.kp, P->F, diva, P->visc_cur, P->visc_oth, P->act_pic.p);
  imau::beta(S.L, S.kp, P->F, diva, P->xg.p, P->yg.p, P->act_pic.p);
  std::swap(P->visc_cur, P->visc_oth);
  imau::coef(S.L, S.kp, P->F, diva, P->act_pic.p);
  S.kcheck();
  std::vector<double> r;
  S.residual(r);
  *residual = r[0];
  return IMAU_OK;
  ABI_CATCH(ctx)
}

int imau_vel_problem_autotune(imau_ctx *ctx, imau_problem *P, const imau_cfg *cfg, int32_t *tile_rows) {
  ABI_TRY
  if (!ctx || !P || !cfg) fail(IMAU_ERR_ARG, "NULL argument");
  check_cfg(cfg);
  if (cfg->nz != P->nz) fail(IMAU_ERR_CONFIG, "cfg nz=%d differs from the problem's nz=%d", cfg->nz, P->nz);
  check_ismip_e(P, cfg);
  const bool diva = cfg->choice_ice_dynamics == IMAU_DYN_DIVA;
  if (!diva || cfg->nz != IMAU_MNZ) {  // only the DIVA viscosity at nz = 15 has a tunable tile
    if (tile_rows) *tile_rows = P->mti;
    return IMAU_OK;
  }
  Solver S(ctx, P, cfg);
  const bool v32 = cfg->precision == IMAU_PREC_MIXED_VISC;
  const long long n3 = P->na * (long long)P->nz;
  if (v32 && !P->visc32.p) P->visc32.alloc((size_t)3 * n3);
  if (v32) {
    imau::to_f32(ctx->s, P->F.B3, P->visc32.p + 2 * n3, n3);
    imau::to_f32(ctx->s, P->visc_cur, P->visc32.p, n3);
  }
  imau::setup1(S.L, S.kp, P->F, P->yg.p);
  imau::fill_i32(ctx->s, P->act_pic.p, P->ngrids, 1);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const int cand[] = IMAU_MTI_CANDIDATES;
  int best = P->mti;
  float best_ms = 0.0f;
  for (int rows : cand) {
    set_mtiles(P, rows);
    auto run = [&]() {
      if (v32) imau::visc3d32(S.L, S.LV(), S.kp, P->F, true, P->visc32.p, P->visc32.p + n3, P->visc32.p + 2 * n3,
                              P->act_pic.p);
      else imau::visc3d(S.L, S.LV(), S.kp, P->F, true, P->visc_cur, P->visc_oth, P->act_pic.p);
    };
    run();  // warm-up
    HIPCHK(hipEventRecord(e0, ctx->s));
    for (int rep = 0; rep < 3; ++rep) run();
    HIPCHK(hipEventRecord(e1, ctx->s));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    if (best_ms == 0.0f || ms < best_ms) { best_ms = ms; best = rows; }
  }
  HIPCHK(hipEventDestroy(e0));
  HIPCHK(hipEventDestroy(e1));
  set_mtiles(P, best);
  HIPCHK(hipGetLastError());
  if (tile_rows) *tile_rows = best;
  return IMAU_OK;
  ABI_CATCH(ctx)
}

int imau_vel_problem_set_tile_rows(imau_ctx *ctx, imau_problem *P, int32_t tile_rows) {
  ABI_TRY
  if (!ctx || !P) fail(IMAU_ERR_ARG, "NULL argument");
  const int cand[] = IMAU_MTI_CANDIDATES;
  bool ok = false;
  for (int rows : cand) ok = ok || rows == tile_rows;
  if (!ok) fail(IMAU_ERR_ARG, "tile_rows=%d is not one of the viscosity kernel's tile heights", (int)tile_rows);
  set_mtiles(P, tile_rows);
  return IMAU_OK;
  ABI_CATCH(ctx)
}

int imau_vel_problem_residual_noalloc(imau_ctx *ctx, imau_problem *P, const imau_cfg *cfg, double *residual_dev) {
  ABI_TRY
  if (!ctx || !P || !cfg || !residual_dev) fail(IMAU_ERR_ARG, "NULL argument");
  check_cfg(cfg);
  if (cfg->nz != P->nz) fail(IMAU_ERR_CONFIG, "cfg nz=%d differs from the problem's nz=%d", cfg->nz, P->nz);
  check_ismip_e(P, cfg);
  const int dyn = cfg->choice_ice_dynamics;
  if (dyn != IMAU_DYN_DIVA && !uses_ssa_field(dyn))
    fail(IMAU_ERR_CONFIG, "the nonlinear residual is defined for DIVA, SSA and SIA/SSA only");
  if (cfg->include_basal_freezing && !P->has_ti) fail(IMAU_ERR_CONFIG, "include_basal_freezing needs Ti_a/Ti_pmp_a at create");
  if (!ctx->scratch) fail(IMAU_ERR_DEVICE, "no scratch address on device %d", ctx->device);
  // layout of the scratch: partials (2 per tile), member sums (2 per member), flags (1 int per member)
  const size_t ntiles = P->tiles.size(), G = (size_t)P->ngrids;
  const size_t need = 2 * ntiles + 2 * G + (G + 1) / 2;
  if (need > kScratchDoubles)
    fail(IMAU_ERR_ARG,
         "problem too large for the no-allocation residual: %zu tiles and %zu members need %zu bytes of scratch, "
         "the maximum is %zu bytes",
         ntiles, G, need * sizeof(double), kScratchDoubles * sizeof(double));
  double *partial = ctx->scratch;
  double *sums = partial + 2 * ntiles;
  int *active = reinterpret_cast<int *>(sums + 2 * G);
  const KP kp = make_kp(cfg);
  const imau::Launch L{ctx->s, (int)ntiles, P->mem_d.p, P->tiles_d.p};
  const imau::Launch LV{ctx->s, P->nvtiles, P->mem_d.p, P->vtiles.p, P->nmtiles, P->mtiles.p};
  const bool diva = dyn == IMAU_DYN_DIVA;
  imau::Coef cf;
  cf.Na = P->F.N_a; cf.Nb = P->F.N_b; cf.Ncx = P->F.N_cx; cf.Ncy = P->F.N_cy;
  cf.bcx = P->F.beta_eff_cx; cf.bcy = P->F.beta_eff_cy; cf.taudx = P->F.taudx; cf.taudy = P->F.taudy;
  cf.frhs = P->F.front_rhs; cf.rt_u = P->F.rt_u; cf.rt_v = P->F.rt_v;
  // the same kernels as imau_vel_nonlinear_residual; the viscosity of the
  // state is kept (the recomputed one goes to the problem's second buffer)
  imau::fill_i32(ctx->s, active, (int)G, 1);
  imau::setup1(L, kp, P->F, P->yg.p);
  imau::setup2(L, kp, P->F);
  imau::visc3d(L, LV, kp, P->F, diva, P->visc_cur, P->visc_oth, active);
  imau::beta(L, kp, P->F, diva, P->xg.p, P->yg.p, active);
  imau::coef(L, kp, P->F, diva, active);
  imau::resid(L, kp, cf, P->F.u, P->F.v, partial);
  imau::member_sum2(ctx->s, (int)G, P->tstart_d.p, partial, sums);
  imau::resid_ratio(ctx->s, (int)G, sums, residual_dev);
  HIPCHK(hipGetLastError());
  return IMAU_OK;
  ABI_CATCH(ctx)
}

}  // extern "C"

// ---------------------------------------------------------------------------
// the ice-thickness update (include/imau_vel.h: imau_thickness_*)
// ---------------------------------------------------------------------------
namespace {

void check_thickness_cfg(const imau_thickness_cfg *c) {
  if (!c) fail(IMAU_ERR_ARG, "thickness cfg is NULL");
  if (c->choice_ice_integration_method < IMAU_DHI_NONE || c->choice_ice_integration_method > IMAU_DHI_SEMIIMPLICIT)
    fail(IMAU_ERR_CONFIG, "unknown choice_ice_integration_method %d", c->choice_ice_integration_method);
  const int *bcs = &c->ice_thickness_west_BC;
  for (int q = 0; q < 4; ++q)
    if (bcs[q] < IMAU_HBC_NONE || bcs[q] > IMAU_HBC_FIXED) fail(IMAU_ERR_CONFIG, "unknown ice_thickness BC %d", bcs[q]);
  if (c->dHi_choice_matrix_solver != IMAU_SOLVER_SOR && c->dHi_choice_matrix_solver != IMAU_SOLVER_PETSC)
    fail(IMAU_ERR_CONFIG, "unknown dHi_choice_matrix_solver %d", c->dHi_choice_matrix_solver);
}

// device buffers of one grid (a-grid-shaped slabs, offset 0)
struct ThkDev {
  DBuf<double> f;    // 8 inputs, 2 outputs, 11 work vectors, 6 rows, x
  DBuf<int32_t> m;   // 4 masks
  DBuf<double> partial, sums;
  imau::ThkIn in{};
  imau::ThkWork w{};
  imau::ThkRows R{};
  double *dHi = nullptr, *Ht = nullptr, *x = nullptr, *sc = nullptr;
  long long na = 0;
  int nblocks = 0;
  void alloc(const Member &mb) {
    na = (long long)mb.nx * mb.ny;
    nblocks = imau::thk_blocks(mb);
    f.alloc((size_t)28 * na);
    m.alloc((size_t)4 * na);
    partial.alloc((size_t)2 * nblocks + 8);
    sums.alloc(16);
    double *q = f.p;
    auto take = [&]() { double *r = q; q += na; return r; };
    double *Hi = take(), *Hb = take(), *SL = take(), *u = take(), *v = take(), *SMB = take(), *BMB = take(),
           *fmf = take();
    in = imau::ThkIn{Hi, Hb, SL, u, v, SMB, BMB, fmf, m.p, m.p + na, m.p + 2 * na, m.p + 3 * na};
    dHi = take();
    Ht = take();
    w.Qx = take(); w.Qy = take(); w.t0 = take(); w.t1 = take();
    w.r = take(); w.rhat = take(); w.p = take(); w.phat = take(); w.v = take(); w.shat = take(); w.t = take();
    R.d = take(); R.cw = take(); R.ce = take(); R.cs = take(); R.cn = take(); R.b = take();
    x = take();
    w.partial = partial.p;
    sc = sums.p + 8;
  }
};

// the update on the device buffers: D.dHi, D.Ht (and the explicit scheme's fluxes)
void thickness_compute(imau_ctx *ctx, const Member &mb, ThkDev &D, const imau_thickness_cfg *c, double dt) {
  hipStream_t s = ctx->s;
  const size_t d8 = sizeof(double);
  const long long na = D.na;
  const int method = c->choice_ice_integration_method;
  if (method == IMAU_DHI_NONE) {
    HIPCHK(hipMemsetAsync(D.dHi, 0, na * d8, s));
    HIPCHK(hipMemcpyAsync(D.Ht, D.in.Hi, na * d8, hipMemcpyDeviceToDevice, s));
  } else if (method == IMAU_DHI_EXPLICIT) {
    imau::thk_explicit(s, mb, D.in, dt, D.w, D.dHi, D.Ht);
  } else {
    imau::thk_rows(s, mb, D.in, dt, D.R);
    HIPCHK(hipMemcpyAsync(D.x, D.in.Hi, na * d8, hipMemcpyDeviceToDevice, s));
    double h[4];
    auto fetch = [&](int n) {
      HIPCHK(hipMemcpyAsync(h, D.sums.p, n * d8, hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
    };
    double scal[4] = {1.0, 1.0, 1.0, 0.0};  // rho_old, alpha, omega, rho
    auto push = [&]() {
      HIPCHK(hipMemcpyAsync(D.sc, scal, sizeof scal, hipMemcpyHostToDevice, s));
      HIPCHK(hipStreamSynchronize(s));  // scal is reused on the host
    };
    imau::thk_r0(s, mb, D.R, D.x, D.w, D.sums.p);
    fetch(2);
    const bool petsc = c->dHi_choice_matrix_solver == IMAU_SOLVER_PETSC;
    const double target = std::max(c->dHi_PETSc_rtol * std::sqrt(h[1]), c->dHi_PETSc_abstol);
    const int maxit = petsc ? 10000 : std::max(1, c->dHi_SOR_nit);
    bool done = petsc && std::sqrt(h[0]) <= target;
    scal[3] = h[0];  // (rhat, r) = |r|^2 at the start
    for (int it = 0; it < maxit && !done; ++it) {
      push();
      imau::thk_bicg_p(s, mb, D.R, D.w, D.sc, D.sums.p);
      fetch(1);
      if (!(std::fabs(h[0]) > 0.0)) break;  // breakdown: the residual is already zero in exact arithmetic
      scal[1] = scal[3] / h[0];
      push();
      imau::thk_bicg_s(s, mb, D.R, D.w, D.sc, D.sums.p);
      fetch(2);
      scal[2] = h[1] > 0.0 ? h[0] / h[1] : 0.0;
      push();
      imau::thk_bicg_x(s, mb, D.x, D.w, D.sc, D.sums.p);
      fetch(3);
      if (!std::isfinite(h[0])) fail(IMAU_ERR_NAN, "non-finite residual in the semi-implicit thickness solve");
      scal[0] = scal[3];
      scal[3] = h[1];
      done = petsc ? std::sqrt(h[0]) <= target : h[2] <= c->dHi_SOR_tol;
      if (scal[2] == 0.0) break;
    }
    imau::thk_finish(s, mb, D.in.Hi, D.x, dt, D.dHi, D.Ht);
  }
  imau::thk_bc(s, mb, D.in, &c->ice_thickness_west_BC, c->choice_ice_dynamics == IMAU_DYN_DIVA, dt, D.dHi, D.Ht);
  HIPCHK(hipGetLastError());
}

// the update on the device buffers, then the download into io (Qx/Qy of the explicit scheme)
void thickness_run(imau_ctx *ctx, const Member &mb, ThkDev &D, const imau_thickness_cfg *c, imau_thickness_io *io,
                   double dt) {
  thickness_compute(ctx, mb, D, c, dt);
  hipStream_t s = ctx->s;
  const size_t d8 = sizeof(double);
  const long long na = D.na;
  const int method = c->choice_ice_integration_method;
  const int nx = mb.nx, ny = mb.ny;
  if (io->dHi_dt_a) HIPCHK(hipMemcpyAsync(io->dHi_dt_a, D.dHi, na * d8, hipMemcpyDeviceToHost, s));
  if (io->Hi_tplusdt_a) HIPCHK(hipMemcpyAsync(io->Hi_tplusdt_a, D.Ht, na * d8, hipMemcpyDeviceToHost, s));
  if (method == IMAU_DHI_EXPLICIT) {
    if (io->Qx_cx) HIPCHK(hipMemcpyAsync(io->Qx_cx, D.w.Qx, (size_t)ny * (nx - 1) * d8, hipMemcpyDeviceToHost, s));
    if (io->Qy_cy)
      HIPCHK(hipMemcpy2DAsync(io->Qy_cy, (size_t)(ny - 1) * d8, D.w.Qy, (size_t)ny * d8, (size_t)(ny - 1) * d8, nx,
                              hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipStreamSynchronize(s));
}

void thickness_upload_extra(imau_ctx *ctx, ThkDev &D, const imau_thickness_io *io, int nx, int ny) {
  hipStream_t s = ctx->s;
  const size_t na = (size_t)nx * ny, d8 = sizeof(double), i4 = sizeof(int32_t);
  if (!io->SMB_year || !io->BMB || !io->float_margin_frac_a || !io->mask_shelf_a || !io->mask_cf_a)
    fail(IMAU_ERR_ARG, "thickness update: SMB_year, BMB, float_margin_frac_a, mask_shelf_a and mask_cf_a are required");
  HIPCHK(hipMemcpyAsync((void *)D.in.SMB, io->SMB_year, na * d8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync((void *)D.in.BMB, io->BMB, na * d8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync((void *)D.in.fmf, io->float_margin_frac_a, na * d8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync((void *)D.in.m_shelf, io->mask_shelf_a, na * i4, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync((void *)D.in.m_cf, io->mask_cf_a, na * i4, hipMemcpyHostToDevice, s));
}

}  // namespace

extern "C" {

void imau_thickness_default_config(imau_thickness_cfg *c) {
  std::memset(c, 0, sizeof *c);
  c->choice_ice_integration_method = IMAU_DHI_EXPLICIT;
  c->choice_ice_dynamics = IMAU_DYN_DIVA;
  c->ice_thickness_west_BC = c->ice_thickness_east_BC = IMAU_HBC_ZERO;
  c->ice_thickness_south_BC = c->ice_thickness_north_BC = IMAU_HBC_ZERO;
  c->dHi_choice_matrix_solver = IMAU_SOLVER_SOR;
  c->dHi_SOR_nit = 3000;
  c->dHi_SOR_tol = 2.5;
  c->dHi_SOR_omega = 1.3;
  c->dHi_PETSc_rtol = 0.001;
  c->dHi_PETSc_abstol = 0.001;
  c->dt_max = 10.0;
}

int imau_thickness_update(imau_ctx *ctx, const imau_grid *grid, const imau_thickness_cfg *cfg, imau_thickness_io *io,
                          double dt) {
  ABI_TRY
  if (!ctx || !grid || !cfg || !io) fail(IMAU_ERR_ARG, "NULL argument");
  check_thickness_cfg(cfg);
  if (grid->nx < 4 || grid->ny < 4) fail(IMAU_ERR_ARG, "grid must be at least 4x4");
  if (!(dt > 0.0)) fail(IMAU_ERR_ARG, "dt must be > 0, got %g", dt);
  if (!io->Hi_a || !io->Hb_a || !io->SL_a || !io->u_vav_cx || !io->v_vav_cy || !io->mask_ice_a || !io->mask_ocean_a)
    fail(IMAU_ERR_ARG, "thickness update: a required input is NULL");
  const int nx = grid->nx, ny = grid->ny;
  const Member mb{nx, ny, grid->dx, 0, 0, 0, 0};
  ThkDev D;
  D.alloc(mb);
  hipStream_t s = ctx->s;
  const size_t na = (size_t)nx * ny, d8 = sizeof(double), i4 = sizeof(int32_t);
  HIPCHK(hipMemcpyAsync((void *)D.in.Hi, io->Hi_a, na * d8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync((void *)D.in.Hb, io->Hb_a, na * d8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync((void *)D.in.SL, io->SL_a, na * d8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync((void *)D.in.u, io->u_vav_cx, (size_t)ny * (nx - 1) * d8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpy2DAsync((void *)D.in.v, (size_t)ny * d8, io->v_vav_cy, (size_t)(ny - 1) * d8, (size_t)(ny - 1) * d8,
                          nx, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync((void *)D.in.m_ice, io->mask_ice_a, na * i4, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync((void *)D.in.m_ocean, io->mask_ocean_a, na * i4, hipMemcpyHostToDevice, s));
  thickness_upload_extra(ctx, D, io, nx, ny);
  thickness_run(ctx, mb, D, cfg, io, dt);
  return IMAU_OK;
  ABI_CATCH(ctx)
}

int imau_vel_problem_thickness(imau_ctx *ctx, imau_problem *P, int32_t g, const imau_thickness_cfg *cfg,
                               imau_thickness_io *io, double dt) {
  ABI_TRY
  if (!ctx || !P || !cfg || !io) fail(IMAU_ERR_ARG, "NULL argument");
  check_thickness_cfg(cfg);
  if (g < 0 || g >= P->ngrids) fail(IMAU_ERR_ARG, "member %d outside [0, %d)", g, P->ngrids);
  if (!(dt > 0.0)) fail(IMAU_ERR_ARG, "dt must be > 0, got %g", dt);
  const Member &pm = P->mem[g];
  const int nx = pm.nx, ny = pm.ny;
  const Member mb{nx, ny, pm.dx, 0, 0, 0, 0};
  ThkDev D;
  D.alloc(mb);
  hipStream_t s = ctx->s;
  const size_t na = (size_t)nx * ny, d8 = sizeof(double), i4 = sizeof(int32_t);
  // the member's slabs, device to device: the solved velocities stay on the GPU
  HIPCHK(hipMemcpyAsync((void *)D.in.Hi, P->F.Hi + pm.off, na * d8, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync((void *)D.in.Hb, P->F.Hb + pm.off, na * d8, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync((void *)D.in.SL, P->F.SL + pm.off, na * d8, hipMemcpyDeviceToDevice, s));
  const double *u = (P->dyn_uploaded == IMAU_DYN_DIVA) ? P->F.u : P->O.uvav;
  const double *v = (P->dyn_uploaded == IMAU_DYN_DIVA) ? P->F.v : P->O.vvav;
  HIPCHK(hipMemcpyAsync((void *)D.in.u, u + pm.off, na * d8, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync((void *)D.in.v, v + pm.off, na * d8, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync((void *)D.in.m_ice, P->F.m_ice + pm.off, na * i4, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync((void *)D.in.m_ocean, P->F.m_ocean + pm.off, na * i4, hipMemcpyDeviceToDevice, s));
  thickness_upload_extra(ctx, D, io, nx, ny);
  thickness_run(ctx, mb, D, cfg, io, dt);
  return IMAU_OK;
  ABI_CATCH(ctx)
}

int imau_critical_timestep_adv(imau_ctx *ctx, const imau_grid *grid, const double *u_vav_cx, const double *v_vav_cy,
                               double dt_max, double *dt_crit_adv) {
  ABI_TRY
  if (!ctx || !grid || !u_vav_cx || !v_vav_cy || !dt_crit_adv) fail(IMAU_ERR_ARG, "NULL argument");
  const int nx = grid->nx, ny = grid->ny;
  const Member mb{nx, ny, grid->dx, 0, 0, 0, 0};
  const size_t na = (size_t)nx * ny, d8 = sizeof(double);
  const int nb = imau::thk_blocks(mb);
  DBuf<double> buf;
  buf.alloc(2 * na + nb + 1);
  hipStream_t s = ctx->s;
  HIPCHK(hipMemcpyAsync(buf.p, u_vav_cx, (size_t)ny * (nx - 1) * d8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpy2DAsync(buf.p + na, (size_t)ny * d8, v_vav_cy, (size_t)(ny - 1) * d8, (size_t)(ny - 1) * d8, nx,
                          hipMemcpyHostToDevice, s));
  imau::cfl(s, mb, buf.p, buf.p + na, buf.p + 2 * na, buf.p + 2 * na + nb);
  HIPCHK(hipGetLastError());
  double m = 0.0;
  HIPCHK(hipMemcpyAsync(&m, buf.p + 2 * na + nb, d8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  *dt_crit_adv = std::min(dt_max, std::min(2.0 * dt_max, m) * 0.9);
  return IMAU_OK;
  ABI_CATCH(ctx)
}

}  // extern "C"

// ---------------------------------------------------------------------------
// the warm-start network (include/imau_net.h)
// ---------------------------------------------------------------------------
struct imau_net_plan {
  int nimg = 0;
  std::vector<int> rows0, cols0;
  long long npix[IMAU_NET_LEVELS] = {0, 0, 0, 0};
  int ntiles[IMAU_NET_LEVELS] = {0, 0, 0, 0};
  DBuf<imau::net::NImg> img[IMAU_NET_LEVELS];
  DBuf<imau::net::NTile> tiles[IMAU_NET_LEVELS];
  DBuf<long long> poff;  // level-0 pixel offset per image
  imau::net::LevelView view(int l) const {
    return imau::net::LevelView{img[l].p, tiles[l].p, ntiles[l], npix[l]};
  }
};

namespace {
void net_plan_build(imau_net_plan *N, int nimg, const int32_t *rows, const int32_t *cols) {
  N->nimg = nimg;
  N->rows0.assign(rows, rows + nimg);
  N->cols0.assign(cols, cols + nimg);
  std::vector<long long> off0(nimg);
  for (int l = 0; l < IMAU_NET_LEVELS; ++l) {
    std::vector<imau::net::NImg> im(nimg);
    std::vector<imau::net::NTile> tl;
    long long off = 0;
    for (int g = 0; g < nimg; ++g) {
      int r = rows[g], c = cols[g];
      for (int k = 0; k < l; ++k) { r = (r + 1) / 2; c = (c + 1) / 2; }
      im[g] = imau::net::NImg{r, c, off};
      if (l == 0) off0[g] = off;
      off += (long long)r * c;
      for (int r0 = 0; r0 < r; r0 += IMAU_NET_TR)
        for (int c0 = 0; c0 < c; c0 += IMAU_NET_TC) tl.push_back(imau::net::NTile{g, r0, c0, 0});
    }
    N->npix[l] = off;
    N->ntiles[l] = (int)tl.size();
    N->img[l].alloc(nimg);
    N->tiles[l].alloc(tl.size());
    HIPCHK(hipMemcpy(N->img[l].p, im.data(), nimg * sizeof(imau::net::NImg), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(N->tiles[l].p, tl.data(), tl.size() * sizeof(imau::net::NTile), hipMemcpyHostToDevice));
  }
  N->poff.alloc(nimg);
  HIPCHK(hipMemcpy(N->poff.p, off0.data(), nimg * sizeof(long long), hipMemcpyHostToDevice));
}

void net_check_level(const imau_net_plan *N, int level, bool coarser) {
  if (!N) fail(IMAU_ERR_ARG, "NULL plan");
  const int top = coarser ? IMAU_NET_LEVELS - 2 : IMAU_NET_LEVELS - 1;
  if (level < 0 || level > top) fail(IMAU_ERR_ARG, "level %d outside [0, %d]", level, top);
}

void net_check_problem(const imau_problem *P, const imau_net_plan *N) {
  if (!P || !N) fail(IMAU_ERR_ARG, "NULL argument");
  if (N->nimg != P->ngrids) fail(IMAU_ERR_ARG, "plan has %d images, problem %d members", N->nimg, P->ngrids);
  for (int g = 0; g < P->ngrids; ++g)
    if (N->rows0[g] != P->mem[g].nx || N->cols0[g] != P->mem[g].ny)
      fail(IMAU_ERR_ARG, "plan image %d is %d x %d, member grid is %d x %d", g, N->rows0[g], N->cols0[g],
           P->mem[g].nx, P->mem[g].ny);
}

bool net_pair_ok(int cin, int cout) {
  return (cin == 16 && cout == 32) || (cin == 32 && cout == 32) || (cin == 32 && cout == 16);
}
}  // namespace

extern "C" {

int imau_net_plan_create(imau_ctx *ctx, int32_t nimg, const int32_t *rows, const int32_t *cols,
                         imau_net_plan **plan) {
  imau_net_plan *N = nullptr;
  ABI_TRY
  if (!ctx || !rows || !cols || !plan) fail(IMAU_ERR_ARG, "NULL argument");
  if (nimg < 1) fail(IMAU_ERR_ARG, "nimg=%d", nimg);
  for (int g = 0; g < nimg; ++g)
    if (rows[g] < 1 || cols[g] < 1) fail(IMAU_ERR_ARG, "image %d is %d x %d", g, rows[g], cols[g]);
  N = new imau_net_plan();
  net_plan_build(N, nimg, rows, cols);
  *plan = N;
  return IMAU_OK;
  }
  catch (const Err &e) {
    if (ctx) ctx->err = e.msg;
    delete N;
    return e.code;
  }
}

int imau_net_plan_for_problem(imau_ctx *ctx, const imau_problem *P, imau_net_plan **plan) {
  ABI_TRY
  if (!ctx || !P || !plan) fail(IMAU_ERR_ARG, "NULL argument");
  std::vector<int32_t> r(P->ngrids), c(P->ngrids);
  for (int g = 0; g < P->ngrids; ++g) { r[g] = P->mem[g].nx; c[g] = P->mem[g].ny; }
  return imau_net_plan_create(ctx, P->ngrids, r.data(), c.data(), plan);
  ABI_CATCH(ctx)
}

void imau_net_plan_destroy(imau_net_plan *plan) { delete plan; }

int64_t imau_net_plan_pixels(const imau_net_plan *plan, int32_t level) {
  if (!plan || level < 0 || level >= IMAU_NET_LEVELS) return -1;
  return plan->npix[level];
}

int64_t imau_net_grad_w_work(const imau_net_plan *plan, int32_t level, int32_t cin, int32_t cout) {
  if (!plan || level < 0 || level >= IMAU_NET_LEVELS || !net_pair_ok(cin, cout)) return -1;
  return (int64_t)imau::net::dw_blocks(plan->view(level)) * (9 * cin * cout + cout);
}

int64_t imau_net_loss_work(const imau_net_plan *plan) {
  if (!plan) return -1;
  return imau::net::loss_blocks(plan->view(0));
}

int imau_net_conv(imau_ctx *ctx, const imau_net_plan *plan, int32_t level, int32_t cin, int32_t cout,
                  const uint16_t *x, const uint16_t *w, const float *bias, int32_t relu, void *y, int32_t y_f32) {
  ABI_TRY
  if (!ctx || !x || !w || !y) fail(IMAU_ERR_ARG, "NULL argument");
  net_check_level(plan, level, false);
  if (!net_pair_ok(cin, cout)) fail(IMAU_ERR_ARG, "unsupported convolution %d -> %d channels", cin, cout);
  if (y_f32 && cout != IMAU_NET_OUT) fail(IMAU_ERR_ARG, "FP32 output only for the %d-channel last layer", IMAU_NET_OUT);
  if (!imau::net::conv(ctx->s, plan->view(level), cin, cout, x, w, bias, relu, y, y_f32))
    fail(IMAU_ERR_ARG, "unsupported convolution %d -> %d channels", cin, cout);
  HIPCHK(hipGetLastError());
  return IMAU_OK;
  ABI_CATCH(ctx)
}

int imau_net_conv_grad_w(imau_ctx *ctx, const imau_net_plan *plan, int32_t level, int32_t cin, int32_t cout,
                         const uint16_t *g, const uint16_t *x, float *dw, float *db, float *work) {
  ABI_TRY
  if (!ctx || !g || !x || !dw || !db || !work) fail(IMAU_ERR_ARG, "NULL argument");
  net_check_level(plan, level, false);
  if (!net_pair_ok(cin, cout)) fail(IMAU_ERR_ARG, "unsupported convolution %d -> %d channels", cin, cout);
  if (!imau::net::conv_dw(ctx->s, plan->view(level), cin, cout, g, x, dw, db, work))
    fail(IMAU_ERR_ARG, "unsupported convolution %d -> %d channels", cin, cout);
  HIPCHK(hipGetLastError());
  return IMAU_OK;
  ABI_CATCH(ctx)
}

int imau_net_pool(imau_ctx *ctx, const imau_net_plan *plan, int32_t level, const uint16_t *x, uint16_t *y) {
  ABI_TRY
  if (!ctx || !x || !y) fail(IMAU_ERR_ARG, "NULL argument");
  net_check_level(plan, level, true);
  imau::net::pool(ctx->s, plan->view(level), plan->view(level + 1), x, y);
  HIPCHK(hipGetLastError());
  return IMAU_OK;
  ABI_CATCH(ctx)
}

int imau_net_upadd(imau_ctx *ctx, const imau_net_plan *plan, int32_t level, const uint16_t *a, const uint16_t *b,
                   uint16_t *y) {
  ABI_TRY
  if (!ctx || !a || !b || !y) fail(IMAU_ERR_ARG, "NULL argument");
  net_check_level(plan, level, true);
  imau::net::upadd(ctx->s, plan->view(level), plan->view(level + 1), a, b, y);
  HIPCHK(hipGetLastError());
  return IMAU_OK;
  ABI_CATCH(ctx)
}

int imau_net_grad_pool(imau_ctx *ctx, const imau_net_plan *plan, int32_t level, const uint16_t *a,
                       const uint16_t *pc, const uint16_t *yact, uint16_t *g) {
  ABI_TRY
  if (!ctx || !yact || !g) fail(IMAU_ERR_ARG, "NULL argument");
  net_check_level(plan, level, pc != nullptr);
  imau::net::gpool(ctx->s, plan->view(level), pc ? plan->view(level + 1) : plan->view(level), a, pc, yact, g);
  HIPCHK(hipGetLastError());
  return IMAU_OK;
  ABI_CATCH(ctx)
}

int imau_net_grad_up(imau_ctx *ctx, const imau_net_plan *plan, int32_t level, const uint16_t *a,
                     const uint16_t *yact, uint16_t *g) {
  ABI_TRY
  if (!ctx || !a || !yact || !g) fail(IMAU_ERR_ARG, "NULL argument");
  net_check_level(plan, level, true);
  imau::net::gup(ctx->s, plan->view(level), plan->view(level + 1), a, yact, g);
  HIPCHK(hipGetLastError());
  return IMAU_OK;
  ABI_CATCH(ctx)
}

int imau_net_loss(imau_ctx *ctx, const imau_net_plan *plan, const float *out, const float *target,
                  const float *weight, float scale, uint16_t *g, double *work, double *loss) {
  ABI_TRY
  if (!ctx || !plan || !out || !target || !weight || !g || !work) fail(IMAU_ERR_ARG, "NULL argument");
  const imau::net::LevelView V = plan->view(0);
  imau::net::loss(ctx->s, V, out, target, weight, scale, g, work);
  HIPCHK(hipGetLastError());
  if (loss) {
    const int nb = imau::net::loss_blocks(V);
    std::vector<double> h(nb);
    HIPCHK(hipMemcpyAsync(h.data(), work, nb * sizeof(double), hipMemcpyDeviceToHost, ctx->s));
    HIPCHK(hipStreamSynchronize(ctx->s));
    double sum = 0.0;
    for (int b = 0; b < nb; ++b) sum += h[b];
    *loss = scale * sum;
  }
  return IMAU_OK;
  ABI_CATCH(ctx)
}

int imau_net_adam(imau_ctx *ctx, int64_t n, float *p, const float *g, float *m, float *v, float lr, float beta1,
                  float beta2, float eps, int32_t step) {
  ABI_TRY
  if (!ctx || !p || !g || !m || !v) fail(IMAU_ERR_ARG, "NULL argument");
  if (n < 0 || step < 1) fail(IMAU_ERR_ARG, "n=%lld step=%d", (long long)n, step);
  const float bc1 = 1.0f - std::pow(beta1, (float)step), bc2 = 1.0f - std::pow(beta2, (float)step);
  imau::net::adam(ctx->s, n, p, g, m, v, lr, beta1, beta2, eps, bc1, bc2);
  HIPCHK(hipGetLastError());
  return IMAU_OK;
  ABI_CATCH(ctx)
}

int imau_net_pack(imau_ctx *ctx, int32_t cin, int32_t cout, const float *w, uint16_t *wf, uint16_t *wt) {
  ABI_TRY
  if (!ctx || !w || !wf || !wt) fail(IMAU_ERR_ARG, "NULL argument");
  if (!net_pair_ok(cin, cout)) fail(IMAU_ERR_ARG, "unsupported convolution %d -> %d channels", cin, cout);
  imau::net::pack(ctx->s, cin, cout, w, wf, wt);
  HIPCHK(hipGetLastError());
  return IMAU_OK;
  ABI_CATCH(ctx)
}

int imau_net_features(imau_ctx *ctx, imau_problem *P, const imau_cfg *cfg, const imau_net_plan *plan,
                      uint16_t *x) {
  ABI_TRY
  if (!ctx || !cfg || !x) fail(IMAU_ERR_ARG, "NULL argument");
  net_check_problem(P, plan);
  const KP kp = make_kp(cfg);
  const imau::Launch L{ctx->s, (int)P->tiles.size(), P->mem_d.p, P->tiles_d.p};
  imau::net_features(L, kp, P->F, plan->poff.p, x);
  HIPCHK(hipGetLastError());
  return IMAU_OK;
  ABI_CATCH(ctx)
}

int imau_net_targets(imau_ctx *ctx, imau_problem *P, const imau_net_plan *plan, float *target, float *weight) {
  ABI_TRY
  if (!ctx || !target || !weight) fail(IMAU_ERR_ARG, "NULL argument");
  net_check_problem(P, plan);
  const imau::Launch L{ctx->s, (int)P->tiles.size(), P->mem_d.p, P->tiles_d.p};
  imau::net_targets(L, P->F, plan->poff.p, target, weight);
  HIPCHK(hipGetLastError());
  return IMAU_OK;
  ABI_CATCH(ctx)
}

int imau_net_warm_start(imau_ctx *ctx, imau_problem *P, const imau_cfg *cfg, const imau_net_plan *plan,
                        const float *out) {
  ABI_TRY
  if (!ctx || !cfg || !out) fail(IMAU_ERR_ARG, "NULL argument");
  net_check_problem(P, plan);
  const KP kp = make_kp(cfg);
  const imau::Launch L{ctx->s, (int)P->tiles.size(), P->mem_d.p, P->tiles_d.p};
  HIPCHK(hipMemcpyAsync(P->act_lin.p, P->sheet_any.data(), P->ngrids * sizeof(int), hipMemcpyHostToDevice, ctx->s));
  imau::net_warm(L, kp, P->F, plan->poff.p, out, P->act_lin.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->s));  // the host copy source is pageable
  return IMAU_OK;
  ABI_CATCH(ctx)
}

int imau_net_sync(imau_ctx *ctx) {
  ABI_TRY
  if (!ctx) fail(IMAU_ERR_ARG, "NULL argument");
  HIPCHK(hipStreamSynchronize(ctx->s));
  return IMAU_OK;
  ABI_CATCH(ctx)
}

}  // extern "C"

// ---------------------------------------------------------------------------
// input producers: update_general_ice_model_data and calc_ice_rheology on
// host arrays (one grid per call)
// ---------------------------------------------------------------------------
namespace {

void check_producer_cfg(const imau_producer_cfg *c) {
  if (c->nz < 2 || c->nz > IMAU_NZ_MAX) fail(IMAU_ERR_CONFIG, "nz=%d outside [2, %d]", c->nz, IMAU_NZ_MAX);
  if (c->choice_ice_rheology < IMAU_RHEO_UNIFORM || c->choice_ice_rheology > IMAU_RHEO_MISMIP_MOD)
    fail(IMAU_ERR_CONFIG, "unknown choice_ice_rheology %d", c->choice_ice_rheology);
  if (c->choice_ice_margin != IMAU_MARGIN_BC && c->choice_ice_margin != IMAU_MARGIN_INFINITE_SLAB)
    fail(IMAU_ERR_CONFIG, "unknown choice_ice_margin %d", c->choice_ice_margin);
}

// device buffers of one grid: 2D fields in a-grid-shaped slabs (offset 0),
// 3D fields in the host's (nz, ny, nx) order
struct ProdDev {
  DBuf<double> f, f3;
  DBuf<int32_t> m;
  DBuf<int> flag;
  imau::ProdIn in{};
  imau::Prod2 o{};
  imau::ProdMasks mk{};
  double *Av = nullptr, *Avcx = nullptr, *Avcy = nullptr;
  double *A3 = nullptr, *A3cx = nullptr, *A3cy = nullptr, *Ti = nullptr;
  long long na = 0;
  int nz = 0;
  void alloc(const Member &mb, int nz_, bool three_d) {
    na = (long long)mb.nx * mb.ny;
    nz = nz_;
    f.alloc((size_t)22 * na);
    m.alloc((size_t)10 * na);
    flag.alloc(1);
    double *q = f.p;
    auto take = [&]() { double *r = q; q += na; return r; };
    double *Hi = take(), *Hb = take(), *SL = take(), *dHi = take(), *dHb = take();
    in = imau::ProdIn{Hi, Hb, SL, dHi, dHb};
    o.Hs = take(); o.TAF = take(); o.dHs_dt = take(); o.Hi_cx = take(); o.Hi_cy = take(); o.Hi_b = take();
    o.dHi_dx = take(); o.dHi_dy = take(); o.dHs_dx = take(); o.dHs_dy = take();
    o.dHs_dx_cx = take(); o.dHs_dy_cx = take(); o.dHs_dx_cy = take(); o.dHs_dy_cy = take();
    Av = take(); Avcx = take(); Avcy = take();
    int32_t *w = m.p;
    auto mt = [&]() { int32_t *r = w; w += na; return r; };
    mk.land = mt(); mk.ocean = mt(); mk.ice = mt(); mk.sheet = mt(); mk.shelf = mt();
    mk.coast = mt(); mk.margin = mt(); mk.gl = mt(); mk.cf = mt(); mk.total = mt();
    if (three_d) {
      f3.alloc((size_t)4 * na * nz);
      A3 = f3.p; A3cx = A3 + na * nz; A3cy = A3cx + na * nz; Ti = A3cy + na * nz;
    }
  }
};

// element strides of a host 3D field with nj points along j
imau::Str3 host3(int nz, int nj) { return imau::Str3{1, (long long)nz * nj, (long long)nz}; }

// host <-> device copies of 2D fields: grid kind 'a', 'x' (cx), 'y' (cy), 'b'
template <class T>
void up2(hipStream_t s, T *dev, const T *host, int nx, int ny, char kind) {
  const int ni = (kind == 'x' || kind == 'b') ? nx - 1 : nx, nj = (kind == 'y' || kind == 'b') ? ny - 1 : ny;
  HIPCHK(hipMemcpy2DAsync(dev, (size_t)ny * sizeof(T), host, (size_t)nj * sizeof(T), (size_t)nj * sizeof(T), ni,
                          hipMemcpyHostToDevice, s));
}
template <class T>
void down2(hipStream_t s, T *host, const T *dev, int nx, int ny, char kind) {
  if (!host) return;
  const int ni = (kind == 'x' || kind == 'b') ? nx - 1 : nx, nj = (kind == 'y' || kind == 'b') ? ny - 1 : ny;
  HIPCHK(hipMemcpy2DAsync(host, (size_t)nj * sizeof(T), dev, (size_t)ny * sizeof(T), (size_t)nj * sizeof(T), ni,
                          hipMemcpyDeviceToHost, s));
}

void download_masks(hipStream_t s, const ProdDev &D, imau_producer_io *io, int nx, int ny) {
  down2(s, io->mask_land_a, D.mk.land, nx, ny, 'a');
  down2(s, io->mask_ocean_a, D.mk.ocean, nx, ny, 'a');
  down2(s, io->mask_ice_a, D.mk.ice, nx, ny, 'a');
  down2(s, io->mask_sheet_a, D.mk.sheet, nx, ny, 'a');
  down2(s, io->mask_shelf_a, D.mk.shelf, nx, ny, 'a');
  down2(s, io->mask_coast_a, D.mk.coast, nx, ny, 'a');
  down2(s, io->mask_margin_a, D.mk.margin, nx, ny, 'a');
  down2(s, io->mask_gl_a, D.mk.gl, nx, ny, 'a');
  down2(s, io->mask_cf_a, D.mk.cf, nx, ny, 'a');
  down2(s, io->mask_a, D.mk.total, nx, ny, 'a');
}

}  // namespace

extern "C" {

void imau_producer_default_config(imau_producer_cfg *c) {
  std::memset(c, 0, sizeof *c);
  const double zeta[15] = {0.00, 0.10, 0.20, 0.30, 0.40, 0.50, 0.60, 0.70, 0.80, 0.90, 0.925, 0.95, 0.975, 0.99, 1.00};
  c->nz = 15;
  for (int k = 0; k < 15; ++k) c->zeta[k] = zeta[k];
  c->choice_ice_margin = IMAU_MARGIN_INFINITE_SLAB;
  c->do_ocean_floodfill = 1;
  c->choice_ice_rheology = IMAU_RHEO_HUYBRECHTS1992;
  c->uniform_flow_factor = 1e-16;
  c->m_enh_sheet = 1.0;
  c->m_enh_shelf = 1.0;
}

int imau_update_general_ice_model_data(imau_ctx *ctx, const imau_grid *grid, const imau_producer_cfg *cfg,
                                       imau_producer_io *io) {
  ABI_TRY
  if (!ctx || !grid || !cfg || !io) fail(IMAU_ERR_ARG, "NULL argument");
  check_producer_cfg(cfg);
  if (grid->nx < 4 || grid->ny < 4) fail(IMAU_ERR_ARG, "grid must be at least 4x4");
  if (!io->Hi_a || !io->Hb_a || !io->SL_a) fail(IMAU_ERR_ARG, "Hi_a, Hb_a and SL_a are required");
  const int nx = grid->nx, ny = grid->ny, nz = cfg->nz;
  const Member mb{nx, ny, grid->dx, 0, 0, 0, 0};
  const bool three_d = io->A_flow_3D_a && (io->A_flow_3D_cx || io->A_flow_3D_cy);
  ProdDev D;
  D.alloc(mb, nz, three_d);
  hipStream_t s = ctx->s;
  up2(s, (double *)D.in.Hi, io->Hi_a, nx, ny, 'a');
  up2(s, (double *)D.in.Hb, io->Hb_a, nx, ny, 'a');
  up2(s, (double *)D.in.SL, io->SL_a, nx, ny, 'a');
  imau::ProdIn in = D.in;
  const bool rates = io->dHi_dt_a && io->dHb_dt_a && io->dHs_dt_a && io->mask_land_a;
  if (rates) {
    up2(s, (double *)in.dHi_dt, io->dHi_dt_a, nx, ny, 'a');
    up2(s, (double *)in.dHb_dt, io->dHb_dt_a, nx, ny, 'a');
    up2(s, D.mk.land, io->mask_land_a, nx, ny, 'a');  // the entry mask dHs/dt reads
  } else {
    in.dHi_dt = in.dHb_dt = nullptr;
  }
  imau::prod_point(s, mb, in, D.o, D.mk, cfg->do_ocean_floodfill ? 1 : 0);
  if (cfg->do_ocean_floodfill) {
    // rounds of tile-local growth until a round changes nothing; every round
    // that changes something adds a cell, so at most na rounds
    int changed = 1;
    for (long long round = 0; changed && round <= D.na; ++round) {
      HIPCHK(hipMemsetAsync(D.flag.p, 0, sizeof(int), s));
      imau::prod_flood_round(s, mb, in, D.mk, D.flag.p);
      HIPCHK(hipMemcpyAsync(&changed, D.flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
    }
    imau::prod_land(s, mb, D.mk);
  }
  imau::prod_masks_maps(s, mb, in, D.o, D.mk);
  if (io->A_flow_vav_a) up2(s, D.Av, io->A_flow_vav_a, nx, ny, 'a');
  if (three_d)
    HIPCHK(hipMemcpyAsync(D.A3, io->A_flow_3D_a, (size_t)D.na * nz * sizeof(double), hipMemcpyHostToDevice, s));
  imau::prod_map_A(s, mb, nz, three_d ? D.A3 : nullptr, host3(nz, ny), D.A3cx, host3(nz, ny), D.A3cy, host3(nz, ny - 1),
                   io->A_flow_vav_a ? D.Av : nullptr, D.Avcx, D.Avcy);
  HIPCHK(hipGetLastError());
  down2(s, io->Hs_a, D.o.Hs, nx, ny, 'a');
  down2(s, io->TAF_a, D.o.TAF, nx, ny, 'a');
  if (rates) down2(s, io->dHs_dt_a, D.o.dHs_dt, nx, ny, 'a');
  download_masks(s, D, io, nx, ny);
  down2(s, io->Hi_cx, D.o.Hi_cx, nx, ny, 'x');
  down2(s, io->Hi_cy, D.o.Hi_cy, nx, ny, 'y');
  down2(s, io->Hi_b, D.o.Hi_b, nx, ny, 'b');
  down2(s, io->dHi_dx_a, D.o.dHi_dx, nx, ny, 'a');
  down2(s, io->dHi_dy_a, D.o.dHi_dy, nx, ny, 'a');
  down2(s, io->dHs_dx_a, D.o.dHs_dx, nx, ny, 'a');
  down2(s, io->dHs_dy_a, D.o.dHs_dy, nx, ny, 'a');
  down2(s, io->dHs_dx_cx, D.o.dHs_dx_cx, nx, ny, 'x');
  down2(s, io->dHs_dy_cx, D.o.dHs_dy_cx, nx, ny, 'x');
  down2(s, io->dHs_dx_cy, D.o.dHs_dx_cy, nx, ny, 'y');
  down2(s, io->dHs_dy_cy, D.o.dHs_dy_cy, nx, ny, 'y');
  if (io->A_flow_vav_a) {
    down2(s, io->A_flow_vav_cx, D.Avcx, nx, ny, 'x');
    down2(s, io->A_flow_vav_cy, D.Avcy, nx, ny, 'y');
  }
  if (three_d) {
    if (io->A_flow_3D_cx)
      HIPCHK(hipMemcpyAsync(io->A_flow_3D_cx, D.A3cx, (size_t)(nx - 1) * ny * nz * sizeof(double),
                            hipMemcpyDeviceToHost, s));
    if (io->A_flow_3D_cy)
      HIPCHK(hipMemcpyAsync(io->A_flow_3D_cy, D.A3cy, (size_t)nx * (ny - 1) * nz * sizeof(double),
                            hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipStreamSynchronize(s));
  return IMAU_OK;
  ABI_CATCH(ctx)
}

int imau_calc_ice_rheology(imau_ctx *ctx, const imau_grid *grid, const imau_producer_cfg *cfg, imau_producer_io *io,
                           double time) {
  ABI_TRY
  if (!ctx || !grid || !cfg || !io) fail(IMAU_ERR_ARG, "NULL argument");
  check_producer_cfg(cfg);
  if (grid->nx < 4 || grid->ny < 4) fail(IMAU_ERR_ARG, "grid must be at least 4x4");
  if (!io->A_flow_3D_a || !io->A_flow_vav_a) fail(IMAU_ERR_ARG, "calc_ice_rheology writes A_flow_3D_a and A_flow_vav_a");
  if (!io->mask_ice_a || !io->mask_sheet_a || !io->mask_shelf_a)
    fail(IMAU_ERR_ARG, "calc_ice_rheology reads mask_ice_a, mask_sheet_a and mask_shelf_a");
  const bool huyb = cfg->choice_ice_rheology == IMAU_RHEO_HUYBRECHTS1992;
  if (huyb && !io->Ti_a) fail(IMAU_ERR_ARG, "Huybrechts1992 reads Ti_a");
  const int nx = grid->nx, ny = grid->ny, nz = cfg->nz;
  const Member mb{nx, ny, grid->dx, 0, 0, 0, 0};
  ProdDev D;
  D.alloc(mb, nz, true);
  hipStream_t s = ctx->s;
  up2(s, D.mk.ice, io->mask_ice_a, nx, ny, 'a');
  up2(s, D.mk.sheet, io->mask_sheet_a, nx, ny, 'a');
  up2(s, D.mk.shelf, io->mask_shelf_a, nx, ny, 'a');
  if (huyb) HIPCHK(hipMemcpyAsync(D.Ti, io->Ti_a, (size_t)D.na * nz * sizeof(double), hipMemcpyHostToDevice, s));
  imau::prod_rheology(s, mb, *cfg, time, D.Ti, host3(nz, ny), D.mk, D.A3, host3(nz, ny), D.Av);
  imau::prod_map_A(s, mb, nz, D.A3, host3(nz, ny), D.A3cx, host3(nz, ny), D.A3cy, host3(nz, ny - 1), D.Av, D.Avcx,
                   D.Avcy);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(io->A_flow_3D_a, D.A3, (size_t)D.na * nz * sizeof(double), hipMemcpyDeviceToHost, s));
  down2(s, io->A_flow_vav_a, D.Av, nx, ny, 'a');
  if (io->A_flow_3D_cx)
    HIPCHK(hipMemcpyAsync(io->A_flow_3D_cx, D.A3cx, (size_t)(nx - 1) * ny * nz * sizeof(double), hipMemcpyDeviceToHost, s));
  if (io->A_flow_3D_cy)
    HIPCHK(hipMemcpyAsync(io->A_flow_3D_cy, D.A3cy, (size_t)nx * (ny - 1) * nz * sizeof(double), hipMemcpyDeviceToHost, s));
  down2(s, io->A_flow_vav_cx, D.Avcx, nx, ny, 'x');
  down2(s, io->A_flow_vav_cy, D.Avcy, nx, ny, 'y');
  HIPCHK(hipStreamSynchronize(s));
  return IMAU_OK;
  ABI_CATCH(ctx)
}

}  // extern "C"

// ---------------------------------------------------------------------------
// time stepping with choice_timestepping = 'direct' on one grid, everything
// device-resident between steps: the velocity problem's own fields are the
// geometry the producers update in place
// ---------------------------------------------------------------------------
namespace {

void check_timestep_cfg(const imau_timestep_cfg *c, const imau_refgeo_io *ref) {
  if (!(c->dt_startup_phase > 0.0)) fail(IMAU_ERR_CONFIG, "dt_startup_phase must be > 0");
  if (!(c->dt_min > 0.0) || c->dt_max < c->dt_min) fail(IMAU_ERR_CONFIG, "need 0 < dt_min <= dt_max");
  if (c->choice_calving_law != IMAU_CALVING_NONE && c->choice_calving_law != IMAU_CALVING_THRESHOLD_THICKNESS)
    fail(IMAU_ERR_CONFIG, "unknown choice_calving_law %d", c->choice_calving_law);
  if (c->remove_shelves_larger_than_PD && !(ref && ref->Hi_PD && ref->Hb_PD))
    fail(IMAU_ERR_ARG, "remove_shelves_larger_than_PD needs ref->Hi_PD and ref->Hb_PD");
  if (c->continental_shelf_calving && !(ref && ref->Hi_GIAeq && ref->Hb_GIAeq))
    fail(IMAU_ERR_ARG, "continental_shelf_calving needs ref->Hi_GIAeq and ref->Hb_GIAeq");
}

struct LoopDev {
  DBuf<double> f;       // SMB, BMB, float margin fraction, dHs/dt, A_vav on cx / cy, zero, Hi_eff_cf
  DBuf<int32_t> m;      // coast, margin, grounding line, calving front, total, shelf keep
  DBuf<int> flag;
  DBuf<double> red;     // reduction partials and the result
  DBuf<double> rg;      // reference geometries of the removal options
  DBuf<int32_t> noice;
  double *SMB, *BMB, *fmf, *dHs_dt, *Avcx, *Avcy, *zero, *heff;
  int32_t *keep;
  int nb = 0;
};

}  // namespace

extern "C" {

namespace {
// the time loop of run_ice_dynamics_direct (pc == nullptr) or _pc
int run_loop(imau_ctx *ctx, const imau_grid *grid, const imau_cfg *cfg, const imau_thickness_cfg *tcfg,
             const imau_producer_cfg *pcfg, const imau_timestep_cfg *scfg, imau_ice_io *io,
             const imau_refgeo_io *ref, const double *SMB_year, const double *BMB, double *Hi_end,
             imau_producer_io *geo, imau_timestep_state *st, imau_pc_state *pc, double t_end) {
  if (!ctx || !grid || !cfg || !tcfg || !pcfg || !scfg || !io || !SMB_year || !BMB || !Hi_end || !st)
    fail(IMAU_ERR_ARG, "NULL argument");
  const int dyn = cfg->choice_ice_dynamics;
  if (!pc && dyn != IMAU_DYN_NONE && dyn != IMAU_DYN_SIA && dyn != IMAU_DYN_SSA && dyn != IMAU_DYN_SIA_SSA)
    fail(IMAU_ERR_CONFIG, "direct time stepping works with 'none', 'SIA', 'SSA' or 'SIA/SSA' only");
  if (pc && (!pc->dHidt_Hn_un || !pc->dHi_dt_a)) fail(IMAU_ERR_ARG, "pc state fields are NULL");
  if (pc && pc->t_next < st->time) fail(IMAU_ERR_CONFIG, "pc t_next %g is before the current time %g", pc->t_next, st->time);
  if (pc && !(scfg->pc_epsilon > 0.0 && scfg->pc_eta_min > 0.0 && pc->pc_eta > 0.0 && pc->pc_eta_prev > 0.0 &&
              pc->dt_crit_ice > 0.0 && pc->dt > 0.0))
    fail(IMAU_ERR_CONFIG, "predictor-corrector: pc_epsilon, pc_eta_min, pc_eta, pc_eta_prev, dt_crit_ice and dt "
                          "must be positive");
  check_cfg(cfg);
  check_thickness_cfg(tcfg);
  check_producer_cfg(pcfg);
  check_timestep_cfg(scfg, ref);
  if (pcfg->nz != cfg->nz) fail(IMAU_ERR_CONFIG, "producer nz=%d differs from the solver's nz=%d", pcfg->nz, cfg->nz);
  imau_cfg vcfg = *cfg;
  if (dyn == IMAU_DYN_NONE) vcfg.choice_ice_dynamics = IMAU_DYN_SIA;  // the problem only carries the geometry
  imau_problem *P = nullptr;
  int rc = imau_vel_problem_create(ctx, 1, grid, &vcfg, io, &P);
  if (rc) return rc;
  struct Release {
    imau_problem *&p;
    ~Release() { imau_vel_problem_destroy(p); }
  } release{P};
  hipStream_t s = ctx->s;
  const Member mb = P->mem[0];
  const int nx = mb.nx, ny = mb.ny, nz = cfg->nz;
  const long long na = (long long)nx * ny;
  const size_t d8 = sizeof(double), i4 = sizeof(int32_t);
  LoopDev X;
  X.f.alloc((size_t)8 * na);
  X.m.alloc((size_t)6 * na);
  X.flag.alloc(1);
  X.nb = imau::thk_blocks(mb);
  X.red.alloc((size_t)X.nb + 1);
  {
    double *q = X.f.p;
    X.SMB = q; X.BMB = q + na; X.fmf = q + 2 * na; X.dHs_dt = q + 3 * na; X.Avcx = q + 4 * na; X.Avcy = q + 5 * na;
    X.zero = q + 6 * na; X.heff = q + 7 * na;
  }
  HIPCHK(hipMemcpyAsync(X.SMB, SMB_year, na * d8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(X.BMB, BMB, na * d8, hipMemcpyHostToDevice, s));
  // the removal rules of update_ice_thickness and their fixed inputs
  imau::Removal rem{};
  rem.shelves = scfg->do_remove_shelves ? 1 : 0;
  rem.min_height = scfg->continental_shelf_min_height;
  if (scfg->remove_shelves_larger_than_PD || scfg->continental_shelf_calving) {
    X.rg.alloc((size_t)4 * na);
    if (scfg->remove_shelves_larger_than_PD) {
      rem.Hi_PD = X.rg.p;
      rem.Hb_PD = X.rg.p + na;
      HIPCHK(hipMemcpyAsync(X.rg.p, ref->Hi_PD, na * d8, hipMemcpyHostToDevice, s));
      HIPCHK(hipMemcpyAsync(X.rg.p + na, ref->Hb_PD, na * d8, hipMemcpyHostToDevice, s));
    }
    if (scfg->continental_shelf_calving) {
      rem.Hi_GIA = X.rg.p + 2 * na;
      rem.Hb_GIA = X.rg.p + 3 * na;
      HIPCHK(hipMemcpyAsync(X.rg.p + 2 * na, ref->Hi_GIAeq, na * d8, hipMemcpyHostToDevice, s));
      HIPCHK(hipMemcpyAsync(X.rg.p + 3 * na, ref->Hb_GIAeq, na * d8, hipMemcpyHostToDevice, s));
    }
  }
  if (ref && ref->mask_noice_a) {
    X.noice.alloc((size_t)na);
    rem.noice = X.noice.p;
    HIPCHK(hipMemcpyAsync(X.noice.p, ref->mask_noice_a, na * i4, hipMemcpyHostToDevice, s));
  }
  const bool removal = rem.noice || rem.shelves || rem.Hi_PD || rem.Hi_GIA;
  const int fixed = (scfg->fixed_shelf_geometry ? 1 : 0) | (scfg->fixed_sheet_geometry ? 2 : 0) |
                    (scfg->fixed_grounding_line ? 4 : 0);
  // the problem's fields are the geometry the producers update
  double *Hi = f2p(P, F_Hi);
  const imau::ProdMasks mk{i2p(P, I_land), i2p(P, I_ocean), i2p(P, I_ice), i2p(P, I_sheet), i2p(P, I_shelf),
                           X.m.p, X.m.p + na, X.m.p + 2 * na, X.m.p + 3 * na, X.m.p + 4 * na};
  X.keep = X.m.p + 5 * na;
  imau::Prod2 o{};
  o.Hs = f2p(P, F_Hs); o.TAF = f2p(P, F_TAF); o.dHs_dt = X.dHs_dt;
  o.Hi_cx = f2p(P, F_Hi_cx); o.Hi_cy = f2p(P, F_Hi_cy); o.Hi_b = f2p(P, F_Hi_b);
  o.dHi_dx = f2p(P, F_dHi_dx_a); o.dHi_dy = f2p(P, F_dHi_dy_a); o.dHs_dx = f2p(P, F_dHs_dx_a);
  o.dHs_dy = f2p(P, F_dHs_dy_a); o.dHs_dx_cx = f2p(P, F_dHs_dx_cx); o.dHs_dy_cx = f2p(P, F_dHs_dy_cx);
  o.dHs_dx_cy = f2p(P, F_dHs_dx_cy); o.dHs_dy_cy = f2p(P, F_dHs_dy_cy);
  const imau::Str3 s3{na, (long long)ny, 1};  // k-plane major
  ThkDev T;
  T.alloc(mb);
  // the current velocities: ice%u_vav_cx (DIVA solves it in place; SIA and SSA compose it)
  const double *uvel = dyn == IMAU_DYN_DIVA ? P->F.u : P->O.uvav, *vvel = dyn == IMAU_DYN_DIVA ? P->F.v : P->O.vvav;
  T.in = imau::ThkIn{Hi, f2p(P, F_Hb), f2p(P, F_SL), uvel, vvel, X.SMB, X.BMB, X.fmf,
                     mk.ice, mk.shelf, mk.ocean, mk.cf};
  // predictor-corrector fields: dHi/dt_n, dHi/dt_n-1, dHi/dt*, Hi_old, Hi_pred, dHi_dt_a
  DBuf<double> pcf;
  double *pHn = nullptr, *pHnm1 = nullptr, *pHstar = nullptr, *pHold = nullptr, *pHpred = nullptr, *pdHi = nullptr;
  if (pc) {
    pcf.alloc((size_t)6 * na);
    pHn = pcf.p; pHnm1 = pcf.p + na; pHstar = pcf.p + 2 * na; pHold = pcf.p + 3 * na; pHpred = pcf.p + 4 * na;
    pdHi = pcf.p + 5 * na;
    HIPCHK(hipMemcpyAsync(pHn, pc->dHidt_Hn_un, na * d8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(pdHi, pc->dHi_dt_a, na * d8, hipMemcpyHostToDevice, s));
  }
  auto flood = [&](auto round) {  // rounds of tile-local growth until one changes nothing
    int changed = 1;
    for (long long r = 0; changed && r <= na; ++r) {
      HIPCHK(hipMemsetAsync(X.flag.p, 0, sizeof(int), s));
      round();
      HIPCHK(hipMemcpyAsync(&changed, X.flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
    }
  };
  // update_general_ice_model_data (with the rates of the last step, or none)
  auto producers = [&](const double *dHi) {
    const imau::ProdIn in{Hi, f2p(P, F_Hb), f2p(P, F_SL), dHi, dHi ? f2p(P, F_dHb_dt) : nullptr};
    imau::prod_point(s, mb, in, o, mk, pcfg->do_ocean_floodfill ? 1 : 0);
    if (pcfg->do_ocean_floodfill) {
      flood([&]() { imau::prod_flood_round(s, mb, in, mk, X.flag.p); });
      imau::prod_land(s, mb, mk);
    }
    imau::prod_masks_maps(s, mb, in, o, mk);
    imau::prod_map_A(s, mb, nz, f3p(P, G_A3), s3, P->Acx3, s3, P->Acy3, s3, f2p(P, F_A_vav), X.Avcx, X.Avcy);
  };
  auto scalar = [&](double *dev) {
    double v = 0.0;
    HIPCHK(hipMemcpyAsync(&v, dev, d8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return v;
  };
  // the derived fields of the state at entry, and its floating-margin fraction
  producers(nullptr);
  imau::prod_float_margin(s, mb, Hi, mk, X.fmf, X.heff);
  imau::Zeta zeta{};
  for (int k = 0; k < nz; ++k) zeta.z[k] = cfg->zeta[k];
  const int32_t nit0 = P->sor_nit_;
  const double tol0 = P->sor_tol_, rtol0 = P->petsc_rtol_, atol0 = P->petsc_abstol_;
  double time = st->time;
  st->nsteps = 0;
  while (time < t_end) {
    // :88-99: tighter velocity solves in the start-up phase (r floored at
    // 1e-3: the reference's r = 0 at the very start asks for unbounded iterations)
    double r = 1.0;
    if (time <= scfg->start_time_of_run + scfg->dt_startup_phase)
      r = std::max(1e-3, 0.01 * 0.99 * (time - scfg->start_time_of_run) / scfg->dt_startup_phase);
    P->sor_nit_ = nit0 * (int32_t)std::ceil(1.0 / r);
    P->sor_tol_ = tol0 * r;
    P->petsc_rtol_ = rtol0 * r;
    P->petsc_abstol_ = atol0 * r;
    // :101-108
    double dt_max;
    if (time <= scfg->start_time_of_run + scfg->dt_startup_phase)
      dt_max = scfg->dt_min + (scfg->dt_max - scfg->dt_min) *
                                  std::pow((time - scfg->start_time_of_run) / scfg->dt_startup_phase, 2);
    else if (time >= scfg->end_time_of_run - scfg->dt_startup_phase)
      dt_max = scfg->dt_min + (scfg->dt_max - scfg->dt_min) *
                                  std::pow((scfg->end_time_of_run - time) / scfg->dt_startup_phase, 2);
    else
      dt_max = scfg->dt_max;
    if (pc && dyn != IMAU_DYN_NONE && time == pc->t_next) {  // :312-415
      imau::cfl(s, mb, uvel, vvel, X.red.p, X.red.p + X.nb);
      HIPCHK(hipGetLastError());
      const double dt_adv = std::min(scfg->dt_max, scalar(X.red.p + X.nb) * 0.9);
      pc->dt_crit_ice_prev = pc->dt_crit_ice;
      const double dt_from_pc = std::pow(scfg->pc_epsilon / pc->pc_eta, scfg->pc_k_I + scfg->pc_k_p) *
                                std::pow(scfg->pc_epsilon / pc->pc_eta_prev, -scfg->pc_k_p) * pc->dt;
      double dc = std::min(std::min(scfg->dt_max, 2.0 * pc->dt_crit_ice_prev), std::min(dt_adv, dt_from_pc));
      dc = std::max(scfg->dt_min, std::max(0.5 * pc->dt_crit_ice_prev, dc));
      dc = std::max(scfg->dt_min, std::min(dt_max, dc));
      pc->dt_crit_ice = dc;
      pc->pc_zeta = dc / pc->dt_crit_ice_prev;
      std::swap(pHn, pHnm1);                                       // dHidt_Hnm1_unm1 = dHidt_Hn_un
      thickness_compute(ctx, mb, T, tcfg, dc);                     // at (H_n, u_n)
      HIPCHK(hipMemcpyAsync(pHn, T.dHi, na * d8, hipMemcpyDeviceToDevice, s));
      imau::pc_pred(s, mb, Hi, pHold, pHpred, pHn, pHnm1, dc, pc->pc_zeta);
      producers(T.dHi);
      if (dyn == IMAU_DYN_SIA || dyn == IMAU_DYN_SIA_SSA)
        if ((rc = solve_kind(ctx, P, cfg, SOLVE_SIA))) return rc;
      if (dyn == IMAU_DYN_SSA || dyn == IMAU_DYN_SIA_SSA)
        if ((rc = solve_kind(ctx, P, cfg, SOLVE_SSA))) return rc;
      if (dyn == IMAU_DYN_DIVA)
        if ((rc = solve_kind(ctx, P, cfg, SOLVE_DIVA))) return rc;
      pc->t_next = time + dc;
      thickness_compute(ctx, mb, T, tcfg, dc);                     // at (H*_n+1, u_n+1)
      HIPCHK(hipMemcpyAsync(pHstar, T.dHi, na * d8, hipMemcpyDeviceToDevice, s));
      HIPCHK(hipMemcpyAsync(Hi, pHold, na * d8, hipMemcpyDeviceToDevice, s));
      producers(T.dHi);
      imau::pc_corr(s, mb, Hi, pHpred, pHn, pHstar, dc, pc->pc_zeta, mk.sheet, mk.gl, scfg->pc_eta_min, pdHi,
                    X.red.p, X.red.p + X.nb);
      HIPCHK(hipGetLastError());
      pc->pc_eta_prev = pc->pc_eta;
      pc->pc_eta = scalar(X.red.p + X.nb);
    }
    if (!pc && (dyn == IMAU_DYN_SIA || dyn == IMAU_DYN_SIA_SSA) && time == st->t_next_SIA) {
      if ((rc = solve_kind(ctx, P, cfg, SOLVE_SIA))) return rc;
      imau::crit_sia(s, mb, nz, zeta, P->O.u3sia, P->O.v3sia, P->O.u3, P->O.v3, o.Hi_cx, o.Hi_cy, o.dHs_dx_cx,
                     o.dHs_dy_cy, mk.sheet, X.red.p, X.red.p + X.nb);
      HIPCHK(hipGetLastError());
      double d = 0.9 * std::min(scfg->dt_max, scalar(X.red.p + X.nb));
      d = std::max(scfg->dt_min, std::min(dt_max, d));
      st->dt_crit_SIA = d;
      st->t_next_SIA = time + d;
    }
    if (!pc && (dyn == IMAU_DYN_SSA || dyn == IMAU_DYN_SIA_SSA) && time == st->t_next_SSA) {
      if ((rc = solve_kind(ctx, P, cfg, SOLVE_SSA))) return rc;
      imau::cfl(s, mb, P->O.uvav, P->O.vvav, X.red.p, X.red.p + X.nb);
      HIPCHK(hipGetLastError());
      double d = std::min(scfg->dt_max, std::min(2.0 * scfg->dt_max, scalar(X.red.p + X.nb)) * 0.9);
      d = std::max(scfg->dt_min, std::min(dt_max, d));
      st->dt_crit_SSA = d;
      st->t_next_SSA = time + d;
    }
    // determine_timesteps_and_actions, :838-856 and :933
    double t_next = std::min(t_end, time + scfg->dt_max);
    if (pc) {
      if (dyn != IMAU_DYN_NONE) t_next = std::min(t_next, pc->t_next);
    } else {
      if (dyn == IMAU_DYN_SIA || dyn == IMAU_DYN_SIA_SSA) t_next = std::min(t_next, st->t_next_SIA);
      if (dyn == IMAU_DYN_SSA || dyn == IMAU_DYN_SIA_SSA) t_next = std::min(t_next, st->t_next_SSA);
    }
    const double dt = t_next - time;
    st->dt = dt;
    if (pc) {
      pc->dt = dt;
      // :257-260 for 'none'; :428 Hi_tplusdt = max(0, Hi + dt dHi_dt_a) (the clip follows)
      if (dyn == IMAU_DYN_NONE) HIPCHK(hipMemsetAsync(pdHi, 0, na * d8, s));
      HIPCHK(hipMemcpyAsync(T.dHi, pdHi, na * d8, hipMemcpyDeviceToDevice, s));
      imau::thk_advance(s, mb, Hi, T.dHi, dt, T.Ht);
    } else if (dyn == IMAU_DYN_NONE) {  // :224-236
      HIPCHK(hipMemsetAsync(T.dHi, 0, na * d8, s));
      HIPCHK(hipMemcpyAsync(T.Ht, Hi, na * d8, hipMemcpyDeviceToDevice, s));
    } else {
      thickness_compute(ctx, mb, T, tcfg, dt);
      imau::thk_advance(s, mb, Hi, T.dHi, dt, T.Ht);
    }
    // update_ice_thickness, :457-645
    if (fixed) imau::prod_fix_geometry(s, mb, fixed, Hi, mk, T.Ht);
    imau::thk_clip(s, mb, T.Ht, Hi);
    producers(nullptr);                                   // determine_masks_ice / _transitions
    imau::prod_float_margin(s, mb, Hi, mk, X.fmf, X.heff);
    if (scfg->choice_calving_law == IMAU_CALVING_THRESHOLD_THICKNESS) {  // :574-588, twice
      imau::prod_calve(s, mb, mk, X.heff, scfg->calving_threshold_thickness, Hi);
      producers(nullptr);
      imau::prod_float_margin(s, mb, Hi, mk, X.fmf, X.heff);
      imau::prod_calve(s, mb, mk, X.heff, scfg->calving_threshold_thickness, Hi);
      producers(nullptr);
    }
    imau::prod_shelf_seed(s, mb, mk, X.keep);             // remove_unconnected_shelves
    flood([&]() { imau::prod_shelf_round(s, mb, mk, X.keep, X.flag.p); });
    imau::prod_shelf_remove(s, mb, mk, X.keep, Hi);
    if (removal) imau::prod_remove_ice(s, mb, rem, f2p(P, F_Hb), f2p(P, F_SL), Hi);
    producers(T.dHi);                                     // update_general_ice_model_data
    HIPCHK(hipGetLastError());
    time += dt;
    ++st->nsteps;
  }
  st->time = time;
  P->sor_nit_ = nit0;
  P->sor_tol_ = tol0;
  P->petsc_rtol_ = rtol0;
  P->petsc_abstol_ = atol0;
  // outputs: velocities through the problem's download, the geometry here
  if (dyn != IMAU_DYN_NONE && (rc = imau_vel_problem_download(ctx, P, io))) return rc;
  HIPCHK(hipMemcpyAsync(Hi_end, Hi, na * d8, hipMemcpyDeviceToHost, s));
  if (geo) {
    down2(s, geo->Hs_a, o.Hs, nx, ny, 'a');
    down2(s, geo->TAF_a, o.TAF, nx, ny, 'a');
    if (st->nsteps > 0) down2(s, geo->dHs_dt_a, o.dHs_dt, nx, ny, 'a');
    const int32_t *msrc[10] = {mk.land, mk.ocean, mk.ice, mk.sheet, mk.shelf, mk.coast, mk.margin, mk.gl, mk.cf,
                               mk.total};
    int32_t *mdst[10] = {geo->mask_land_a, geo->mask_ocean_a, geo->mask_ice_a, geo->mask_sheet_a, geo->mask_shelf_a,
                         geo->mask_coast_a, geo->mask_margin_a, geo->mask_gl_a, geo->mask_cf_a, geo->mask_a};
    for (int q = 0; q < 10; ++q) down2(s, mdst[q], msrc[q], nx, ny, 'a');
    down2(s, geo->Hi_cx, o.Hi_cx, nx, ny, 'x');
    down2(s, geo->Hi_cy, o.Hi_cy, nx, ny, 'y');
    down2(s, geo->Hi_b, o.Hi_b, nx, ny, 'b');
    down2(s, geo->dHi_dx_a, o.dHi_dx, nx, ny, 'a');
    down2(s, geo->dHi_dy_a, o.dHi_dy, nx, ny, 'a');
    down2(s, geo->dHs_dx_a, o.dHs_dx, nx, ny, 'a');
    down2(s, geo->dHs_dy_a, o.dHs_dy, nx, ny, 'a');
    down2(s, geo->dHs_dx_cx, o.dHs_dx_cx, nx, ny, 'x');
    down2(s, geo->dHs_dy_cx, o.dHs_dy_cx, nx, ny, 'x');
    down2(s, geo->dHs_dx_cy, o.dHs_dx_cy, nx, ny, 'y');
    down2(s, geo->dHs_dy_cy, o.dHs_dy_cy, nx, ny, 'y');
    down2(s, geo->float_margin_frac_a, X.fmf, nx, ny, 'a');
    down2(s, geo->Hi_eff_cf_a, X.heff, nx, ny, 'a');
    down2(s, geo->A_flow_vav_a, f2p(P, F_A_vav), nx, ny, 'a');
    down2(s, geo->A_flow_vav_cx, X.Avcx, nx, ny, 'x');
    down2(s, geo->A_flow_vav_cy, X.Avcy, nx, ny, 'y');
    down3(ctx, P, 0, KA, geo->A_flow_3D_a, f3p(P, G_A3));
    down3(ctx, P, 0, KCX, geo->A_flow_3D_cx, P->Acx3);
    down3(ctx, P, 0, KCY, geo->A_flow_3D_cy, P->Acy3);
  }
  if (pc) {
    HIPCHK(hipMemcpyAsync(pc->dHidt_Hn_un, pHn, na * d8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(pc->dHi_dt_a, pdHi, na * d8, hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipStreamSynchronize(s));
  (void)i4;
  return IMAU_OK;
}
}  // namespace

int imau_run_ice_dynamics_direct(imau_ctx *ctx, const imau_grid *grid, const imau_cfg *cfg,
                                 const imau_thickness_cfg *tcfg, const imau_producer_cfg *pcfg,
                                 const imau_timestep_cfg *scfg, imau_ice_io *io, const imau_refgeo_io *ref,
                                 const double *SMB_year, const double *BMB, double *Hi_end, imau_producer_io *geo,
                                 imau_timestep_state *st, double t_end) {
  ABI_TRY
  return run_loop(ctx, grid, cfg, tcfg, pcfg, scfg, io, ref, SMB_year, BMB, Hi_end, geo, st, nullptr, t_end);
  ABI_CATCH(ctx)
}

int imau_run_ice_dynamics_pc(imau_ctx *ctx, const imau_grid *grid, const imau_cfg *cfg,
                             const imau_thickness_cfg *tcfg, const imau_producer_cfg *pcfg,
                             const imau_timestep_cfg *scfg, imau_ice_io *io, const imau_refgeo_io *ref,
                             const double *SMB_year, const double *BMB, double *Hi_end, imau_producer_io *geo,
                             imau_timestep_state *st, imau_pc_state *pc, double t_end) {
  ABI_TRY
  if (!pc) fail(IMAU_ERR_ARG, "NULL argument");
  return run_loop(ctx, grid, cfg, tcfg, pcfg, scfg, io, ref, SMB_year, BMB, Hi_end, geo, st, pc, t_end);
  ABI_CATCH(ctx)
}

}  // extern "C"
