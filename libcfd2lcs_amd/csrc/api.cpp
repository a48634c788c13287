// api.cpp -- the C ABI of the MI355X-native libcfd2lcs FTLE path and the
// per-update orchestration of the HIP kernels.
//
// Library-global state with a sticky error flag, as the reference keeps
// (data_m.f90:313-381, cfd2lcs.f90:684-706).  One process drives one GPU;
// the MPI communicator is accepted and ignored.  Compiled twice: with
// LCS_SINGLE defined the user-facing real type (lcsdata_t) is float
// (libcfd2lcs_sp.so), otherwise double (libcfd2lcs_dp.so).  The device compute
// precision is a separate run-time option ("PRECISION"), defaulting to the
// library's type.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#ifdef LCS_SINGLE
#define CFD2LCS_LCSDATA_DEFINED
typedef float lcsdata_t;
#else
#define CFD2LCS_LCSDATA_DEFINED
typedef double lcsdata_t;
#endif
#define CFD2LCS_NO_MPI
#include "../../include/cfd2lcs_api.h"
#include "grid_host.h"
#include "lcs_device.h"
#include "lcs_io.h"
#include "scratch.h"

using U = lcsdata_t;  // user-facing real type

namespace lcs {

// largest tile-store scratch of the adjoint scatter (bytes)
constexpr size_t kTileStoreMax = size_t(16) << 30;

// ---------------------------------------------------------------------------
// options (data_m.f90:15-94) and sticky error
// ---------------------------------------------------------------------------
struct Options {
  int integrator = kRK2;
  int interpolator = kLinear;
  int incompressible = 0;
  int verbose = 0;
  int sync_timer = 0;
  U cfl_max = (U)0.5;
  int precision = sizeof(U) == 8 ? LCS_FP64 : LCS_FP32;
  int kernel_timing = 0;  // time every hot kernel launch with HIP events
  int particle_sort = 1;  // re-order forward particles by node every n updates (0: never)
  // composition: one launch per slot with the particles re-ordered by brick
  // before every n-th slot (0: the single-launch composition)
  int compose_sort = 1;
  // B-spline steps stage the basis tables of their workgroup's intervals in
  // LDS (0: every table read from global memory; same results)
  int spline_lds = 1;
  int vel_storage = kStoreNative;  // storage of the gathered stage velocities
  // storage of the backward semi-Lagrangian flowmap and its coefficients
  // (kStoreF32: FP32 under FP64 compute); takes effect at the next h-interval
  // reset
  int fm_storage = kStoreNative;
  int compensated = 0;    // two-sum position/displacement updates of forward particles
  int sl_tiled = 1;       // LDS bricks for the FP32 cubic semi-Lagrangian step (results identical)
  // output files: FLOWMAP_IO, VELOCITY_IO, BCFLAG_IO (data_m.f90:44-54) pick
  // the variables of the Tecplot FTLE file; write_lcs switches that file
  // (always written by the reference) and write_slots the flowmap/velocity
  // slot files that the reference keeps in ./cfd2lcs_tmp (here the slots
  // live in HBM, so the files are optional)
  int flowmap_io = 1, velocity_io = 1, bcflag_io = 0;
  int invariants = 0;  // VELOCITY_INVARIANTS (data_m.f90:59): Q, C, D, H, L2 columns
  int write_lcs = 1;
  int write_slots = 0;
  // AUX_GRID (data_m.f90:39): six auxiliary particle sets per FTLE
  // diagnostic, read when the diagnostic is created
  int aux_grid = 0;
  // adjoint scatter by in-place colour launches instead of the tile store
  // (which it falls back to anyway above kTileStoreMax): for tests
  int adj_colours = 0;
  bool warn_opt = false, warn_par = false;  // SAVEd warn flags (cfd2lcs.f90:444,632)
};

static Options g_opt;
static int g_error = 0;

static void fail(const std::string& msg) {
  std::printf("ERROR: libcfd2lcs: %s\n", msg.c_str());
  std::fflush(stdout);
  g_error = 1;
}

#define HIPCHK(expr)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      fail(std::string("HIP error ") + hipGetErrorString(e_) + " at " #expr);           \
      return;                                                                           \
    }                                                                                   \
  } while (0)
#define HIPCHK_R(expr, rv)                                                              \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      fail(std::string("HIP error ") + hipGetErrorString(e_) + " at " #expr);           \
      return rv;                                                                        \
    }                                                                                   \
  } while (0)

// device allocation owned by the engine, or (owned = false) a view of
// memory that someone else owns (the fixed workspace of scratch.hip)
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool owned = true;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), owned(o.owned) {
    o.p = nullptr;
    o.bytes = 0;
    o.owned = true;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p; bytes = o.bytes; owned = o.owned;
      o.p = nullptr; o.bytes = 0; o.owned = true;
    }
    return *this;
  }
  void swap(DevBuf& o) noexcept {
    std::swap(p, o.p);
    std::swap(bytes, o.bytes);
    std::swap(owned, o.owned);
  }
  ~DevBuf() { release(); }
  void release() {
    if (p && owned) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    owned = true;
  }
  // at least b bytes (the contents are lost when it grows)
  bool ensure(size_t b) { return (p && bytes >= b) || alloc(b); }
  bool alloc(size_t b) {
    release();
    if (b == 0) return true;
    if (hipMalloc(&p, b) != hipSuccess) {
      p = nullptr;
      return false;
    }
    bytes = b;
    return true;
  }
  void view(void* q, size_t b) {
    release();
    p = q;
    bytes = b;
    owned = false;
  }
};

struct EngineBase {
  virtual ~EngineBase() {}
  virtual int add_diag(int type, int resolution, U T, U h, const char* label) = 0;
  virtual void update(const U* u, const U* v, const U* w, bool device, U time) = 0;
  virtual int get_flowmap(int h, U* a, U* b, U* c) = 0;
  virtual int get_composed(int h, U* a, U* b, U* c) = 0;
  virtual int get_ftle(int h, U* out) = 0;
  virtual int get_grid(int h, int n[3], U* x, U* y, U* z) = 0;
  virtual const double* ftle_device(int h) = 0;
  virtual int noutput(int h) = 0;
  virtual void stats(double out[8]) = 0;
  virtual int debug_grid(double* lines, int* flag, double* bn, double* ps, double* bb) = 0;
  virtual bool fresh() const = 0;
  virtual int timers(double* sec, long long* cnt, int n) = 0;
  virtual int ndiag() const = 0;
  virtual int adjoint_forward(int nsnap, const U* du, const U* dv, const U* dw, const U* times, int ckpt,
                              double* d_ftle) = 0;
  virtual int adjoint_backward(const double* d_dftle, U* gu, U* gv, U* gw) = 0;
  virtual int adjoint_forward_noalloc(int nsnap, const U* du, const U* dv, const U* dw, const U* times,
                                      double* d_ftle) = 0;
  virtual void adjoint_release() = 0;
  virtual void adjoint_stats(double out[4]) = 0;
  virtual int flush_output() = 0;
};

// ---------------------------------------------------------------------------
// Engine<T>: device state and orchestration in compute precision T
// ---------------------------------------------------------------------------
template <typename T>
struct Engine : EngineBase {
  HostGrid hg;
  GridDev<T> g{};
  hipStream_t stream = nullptr;
  DevBuf grid_mem;      // lines, flags, normals, deltas
  DevBuf spline_mem[3]; // knots + band per axis
  int spline_k = 0;     // order of the fitted spline state (0 = none)
  DevBuf status_mem;
  DevStatus* st = nullptr;
  bool ws_ok = false;  // the fixed workspace's address was resolved at init
  DevStatus* st_host = nullptr;
  DevBuf vel[2];        // [u_n, u_np1] ghosted, 3 components each
  DevBuf vcoef[2];      // spline coefficients of vel[]
  bool vcoef_ok[2] = {false, false};
  // vcoef[b] holds FP32 coefficients (FP64 compute with FP32 stage storage:
  // the blends then read half the bytes, and ct = 0 or 1 reads them as they
  // are)
  bool vcoef32[2] = {false, false};
  bool want_vcoef32() const { return sizeof(T) == 8 && g_opt.vel_storage == kStoreF32; }
  int cur = 1;          // index of u_np1 in vel[]
  DevBuf stage_u;       // staging of user velocity (lcsdata_t), 3*np
  DevBuf stage_out;     // staging for getters (lcsdata_t), 3*np
  bool first_call = true;
  U t_n = 0, t_np1 = 0;
  int nsub_last = 0;
  long long integ_fwd = 0, integ_bkwd = 0;
  // timers: 0 FWD advect, 1 BKWD advect, 2 composition, 3 FTLE (phases);
  // with KERNEL_TIMING: 8 k_fwd_step, 9 k_sl_step, 10 spline fit,
  // 11 k_compose, 12 k_ftle, 13 ghost/ingest/CFL kernels, 14 time blend,
  // 15 forward particle sort, 16 composition sort
  enum { kTimers = 17 };
  double tsum[kTimers] = {};
  long long tcount[kTimers] = {};
  std::vector<hipEvent_t> evpool;
  unsigned long long oob_total = 0;

  struct Timed {
    hipEvent_t a, b;
    int cat;
  };
  std::vector<Timed> pending;

  // The grid of a particle set other than the flow grid: a diagnostic with
  // resolution != 0 (new_sgrid_from_sgrid) or an AUX_GRID set
  // (create_aux_grid).  Own spline factors; per node the tracked flow node
  // of a backward set (lp%no_scfd after init_lp / reset_lp + track_lp2node).
  struct SetGrid {
    HostGrid hg;
    GridDev<T> g{};
    DevBuf mem, spline_mem[3];
    int spline_k = 0;
    DevBuf nodes;  // 3 x np int16
    std::shared_ptr<std::vector<U>> out_xyz, out_flag;  // output-file columns
  };

  struct Diag {
    int id, type, dir;
    U Tint, h, dt_factor;
    std::string label;
    std::string lp_label;  // lp%label: <label>-LP, AUX_GRID sets <label>-LP0 .. -LP5
    std::unique_ptr<SetGrid> sg;  // null: the particles live on the flow grid
    // with sg, backward: the tracked nodes are unclamped until the first
    // sub-cycle after a reset (set_lp_bc clamps lp%no_scfd in place)
    bool fresh_nodes = true;
    std::vector<std::unique_ptr<Diag>> aux;  // AUX_GRID sets X0, Y0, Z0, X1, Y1, Z1
    int nstep, nslot;
    std::vector<char> slot_ok, slot_coef_ok;
    DevBuf part[2];        // FWD particle sets (current = part[pcur])
    int pcur = 0;
    bool sorted = false;   // part[pcur] is not in origin order
    int since_sort = 0;    // updates advected since the last reset or sort
    int cmode = 0;         // compensated updates in the current h interval
    DevBuf skeys, svals, stemp;  // sort scratch
    size_t stemp_bytes = 0;
    DevBuf fm[2];          // BKWD semi-Lagrangian flowmap ping-pong (ghosted)
    int fcur = 0;
    bool fm32 = false;     // fm and fmcoef hold FP32 in this h interval (FP64 compute)
    size_t fm_elem() const { return fm32 ? sizeof(float) : sizeof(T); }
    DevBuf fmcoef;         // spline coefficients of fm[fcur]
    DevBuf slots, slotcoef;  // nslot * 3 ghosted fields
    DevBuf order;          // slot order of the composition
    std::vector<int> order_h;  // its host copy (kept until the copy ran)
    DevBuf dxout;          // composed displacement by origin (3*np)
    DevBuf compfm;         // composed flowmap, ghosted
    DevBuf comp;           // composed flowmap interior (3*np)
    DevBuf ftle;           // np doubles
    int noutput = 0;
  };
  std::vector<std::unique_ptr<Diag>> diags;

  // a diagnostic's own set, then its AUX_GRID sets (the order of lp_c)
  template <class F>
  void for_sets(Diag& dg, F&& f) {
    f(dg);
    for (auto& a : dg.aux) f(*a);
  }
  GridDev<T>& G(Diag& dg) { return dg.sg ? dg.sg->g : g; }
  const HostGrid& HG(Diag& dg) { return dg.sg ? dg.sg->hg : hg; }
  long long NP(Diag& dg) { return G(dg).np; }
  long long NT(Diag& dg) { return G(dg).nt; }

  // output files (write_lcs, write_flowmap_substep)
  std::unique_ptr<io::Writer> writer;
  std::shared_ptr<std::vector<U>> out_xyz;   // interior coordinates, 3 * np
  std::shared_ptr<std::vector<U>> out_flag;  // interior bcflag as reals, np
  // velocity at h boundaries by slot step, shared by all diagnostics as the
  // reference's ./cfd2lcs_tmp/<rank>_U_NP1_<step>.dat files are
  // (flowmap_m.f90:138-151), read back for the forward FTLE's U columns
  // (flowmap_m.f90:212-247)
  std::vector<DevBuf> uring;
  std::vector<char> uring_ok;
  DevBuf vint;  // interior scratch, 3 * np
  DevBuf iscratch;  // interior scratch of any grid (host copies)
  DevBuf invbuf;  // velocity invariants, 5 * np
  DevBuf uscratch[2];  // velocity at T0 of a diagnostic grid: ghosted field on
                       // the flow grid (and its spline fit), then 3 * np
  // the re-ordered composition's particle records, brick keys / positions
  // and sort scratch, shared by all sets (sized for the largest)
  DevBuf crec[2], ckeys, cvals, ctemp;

  ~Engine() override {
    if (writer) writer->flush();
    if (stream) {
      (void)hipStreamSynchronize(stream);
      for (auto& t : pending) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
      }
      for (auto e : evpool) (void)hipEventDestroy(e);
      (void)hipStreamDestroy(stream);
    }
    if (st_host) (void)hipHostFree(st_host);
  }

  bool fresh() const override { return diags.empty() && first_call; }
  int ndiag() const override { return (int)diags.size(); }
  int timers(double* sec, long long* cnt, int n) override {
    sync_status();
    const int m = std::min(n, (int)kTimers);
    for (int a = 0; a < m; a++) {
      if (sec) sec[a] = tsum[a];
      if (cnt) cnt[a] = tcount[a];
    }
    return m;
  }

  // ghosted 3-component fields of nt values per component (default: the
  // flow grid's)
  size_t fbytes(long long nt) const { return sizeof(T) * (size_t)nt; }
  size_t fbytes() const { return fbytes(g.nt); }
  F3<T> f3(DevBuf& b, long long nt) {
    T* p = (T*)b.p;
    return F3<T>{{p, p + nt, p + 2 * nt}};
  }
  CF3<T> cf3(DevBuf& b, long long nt) {
    const T* p = (const T*)b.p;
    return CF3<T>{{p, p + nt, p + 2 * nt}};
  }
  F3<T> f3(DevBuf& b) { return f3(b, g.nt); }
  CF3<T> cf3(DevBuf& b) { return cf3(b, g.nt); }
  // the three component arrays of b as element type E
  template <typename E>
  static F3<E> f3e(DevBuf& b, long long nt) {
    E* p = (E*)b.p;
    return F3<E>{{p, p + nt, p + 2 * nt}};
  }
  template <typename E>
  static CF3<E> cf3e(DevBuf& b, long long nt) {
    const E* p = (const E*)b.p;
    return CF3<E>{{p, p + nt, p + 2 * nt}};
  }
  // particle-set layout inside one allocation: x[3], d[3] (T), n[3]
  // (int16), org (int), each array aligned to 256 bytes
  size_t pset_bytes(long long np) const {
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    return 12 * al(sizeof(T) * (size_t)np) + 3 * al(sizeof(int16_t) * (size_t)np) + al(sizeof(int) * (size_t)np);
  }
  ParticleSet<T> pset(DevBuf& b, long long np) {
    char* base = (char*)b.p;
    size_t off = 0;
    auto take = [&](size_t bytes) {
      char* q = base + off;
      off += (bytes + 255) & ~size_t(255);
      return q;
    };
    ParticleSet<T> ps;
    for (int d = 0; d < 3; d++) ps.x[d] = (T*)take(sizeof(T) * (size_t)np);
    for (int d = 0; d < 3; d++) ps.d[d] = (T*)take(sizeof(T) * (size_t)np);
    for (int d = 0; d < 3; d++) ps.xl[d] = (T*)take(sizeof(T) * (size_t)np);
    for (int d = 0; d < 3; d++) ps.dl[d] = (T*)take(sizeof(T) * (size_t)np);
    for (int d = 0; d < 3; d++) ps.n[d] = (int16_t*)take(sizeof(int16_t) * (size_t)np);
    ps.org = (int*)take(sizeof(int) * (size_t)np);
    return ps;
  }
  // forward displacement by origin node: the current set's dx, or (when the
  // set is sorted, or dx is a compensated sum) its scatter to origin order in
  // the spare set's dx arrays
  CF3<T> fwd_dx_by_origin(Diag& dg) {
    const long long n = NP(dg);
    ParticleSet<T> a = pset(dg.part[dg.pcur], n);
    if (!dg.sorted && !dg.cmode) return CF3<T>{{a.d[0], a.d[1], a.d[2]}};
    if (!dg.part[1].p && !dg.part[1].alloc(pset_bytes(n))) {
      fail("hipMalloc failed (particle set)");
      return CF3<T>{{a.d[0], a.d[1], a.d[2]}};
    }
    ParticleSet<T> b = pset(dg.part[1 - dg.pcur], n);
    const T* lo[3] = {a.dl[0], a.dl[1], a.dl[2]};
    launch_unpermute(n, a.org, CF3<T>{{a.d[0], a.d[1], a.d[2]}}, dg.cmode ? lo : nullptr,
                     F3<T>{{b.d[0], b.d[1], b.d[2]}}, stream);
    return CF3<T>{{b.d[0], b.d[1], b.d[2]}};
  }
  bool sort_particles(Diag& dg) {
    const long long n = NP(dg);
    const size_t np = (size_t)n;
    if (!dg.stemp.p) {
      dg.stemp_bytes = particle_sort_temp_bytes(n, g.np);
      if (!dg.stemp_bytes || (!dg.part[1].p && !dg.part[1].alloc(pset_bytes(n))) ||
          !dg.skeys.alloc(2 * sizeof(unsigned) * np) ||
          !dg.svals.alloc(2 * sizeof(unsigned) * np) || !dg.stemp.alloc(dg.stemp_bytes)) {
        fail("hipMalloc failed (particle sort buffers)");
        return false;
      }
    }
    if (launch_particle_sort(g, pset(dg.part[dg.pcur], n), pset(dg.part[1 - dg.pcur], n), n, (unsigned*)dg.skeys.p,
                             (unsigned*)dg.svals.p, dg.stemp.p, dg.stemp_bytes, stream)) {
      fail("particle sort failed");
      return false;
    }
    dg.pcur = 1 - dg.pcur;
    dg.sorted = true;
    dg.since_sort = 0;
    return true;
  }
  F3<T> p3(DevBuf& b, long long np) { return f3(b, np); }
  CF3<T> cp3(DevBuf& b, long long np) { return cf3(b, np); }
  F3<T> p3(DevBuf& b) { return p3(b, g.np); }
  CF3<T> cp3(DevBuf& b) { return cp3(b, g.np); }

  // device copy of a grid: coordinate lines, interpolation lines, normals,
  // 1/delta and flags in one allocation
  bool upload_grid(const HostGrid& h, GridDev<T>& gd, DevBuf& mem) {
    for (int d = 0; d < 3; d++) {
      gd.n[d] = h.n[d];
      gd.N[d] = h.N[d];
      gd.per[d] = h.per[d];
      gd.kord[d] = 0;
      gd.kn[d] = nullptr;
      gd.band[d] = nullptr;
      gd.bpoly[d] = nullptr;
    }
    gd.nt = h.nt;
    gd.np = h.np;
    gd.flat = h.flat;
    gd.max_track = 2 * std::max(h.n[0], std::max(h.n[1], h.n[2]));
    for (int c = 0; c < 27; c++)
      for (int d = 0; d < 3; d++) gd.ps[c][d] = (T)h.ps[c][d];
    gd.ps_sep = 1;
    for (int k = -1; k <= 1; k++)
      for (int j = -1; j <= 1; j++)
        for (int i = -1; i <= 1; i++) {
          const int c = (i + 1) + 3 * (j + 1) + 9 * (k + 1);
          const int a[3] = {(i + 1), 3 * (j + 1), 9 * (k + 1)};
          for (int d = 0; d < 3; d++)
            if (gd.ps[c][d] != gd.ps[a[d] + (d == 0 ? 3 + 9 : (d == 1 ? 1 + 9 : 1 + 3))][d]) gd.ps_sep = 0;
        }
    for (int d = 0; d < 6; d++) gd.bb[d] = (T)h.bb[d];
    const size_t nl = (size_t)h.N[0] + h.N[1] + h.N[2];
    // a non-rectilinear grid adds its full coordinates and face weights
    const size_t nr = h.rect ? 0 : 3 * (size_t)h.nt + 18 * (size_t)h.np;
    const size_t reals = 2 * nl + 6 * (size_t)h.np + nr;
    const size_t bytes = reals * sizeof(T) + (size_t)h.nt + 256;
    if (!mem.alloc(bytes)) {
      fail("hipMalloc failed for the grid");
      return false;
    }
    std::vector<T> hr(reals);
    size_t o = 0;
    size_t off_gl[3], off_xl[3], off_bn[3], off_dl[3];
    for (int d = 0; d < 3; d++) {
      off_gl[d] = o;
      for (int a = 0; a < h.N[d]; a++) hr[o++] = (T)h.line[d][a];
      off_xl[d] = o;
      for (int a = 0; a < h.N[d]; a++) hr[o++] = (T)h.xl[d][a];
    }
    for (int d = 0; d < 3; d++) {
      off_bn[d] = o;
      for (long long p = 0; p < h.np; p++) hr[o++] = (T)h.bn[d][p];
      off_dl[d] = o;
      for (long long p = 0; p < h.np; p++) hr[o++] = (T)h.dl[d][p];
    }
    size_t off_gx[3] = {0, 0, 0}, off_lsw = 0;
    if (!h.rect) {
      for (int d = 0; d < 3; d++) {
        off_gx[d] = o;
        for (long long l = 0; l < h.nt; l++) hr[o++] = (T)h.g[d][l];
      }
      off_lsw = o;
      for (size_t a = 0; a < h.lsw.size(); a++) hr[o++] = (T)h.lsw[a];
    }
    T* base = (T*)mem.p;
    if (hipMemcpy(base, hr.data(), reals * sizeof(T), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy((char*)(base + reals), h.flag.data(), (size_t)h.nt, hipMemcpyHostToDevice) != hipSuccess) {
      fail("upload of the grid failed");
      return false;
    }
    for (int d = 0; d < 3; d++) {
      gd.gl[d] = base + off_gl[d];
      gd.xl[d] = base + off_xl[d];
      gd.bn[d] = base + off_bn[d];
      gd.dl[d] = base + off_dl[d];
    }
    gd.flag = (const int8_t*)(base + reals);
    gd.rect = h.rect ? 1 : 0;
    for (int d = 0; d < 3; d++) gd.gx[d] = h.rect ? nullptr : base + off_gx[d];
    gd.lsw = h.rect ? nullptr : base + off_lsw;
    for (int d = 0; d < 3; d++)
      for (int e = 0; e < 3; e++) {
        gd.pmin[d][e] = (T)h.pmin[d][e];
        gd.pmax[d][e] = (T)h.pmax[d][e];
      }
    return true;
  }

  bool init(const HostGrid& h) {
    hg = h;
    if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) {
      fail("cannot create a HIP stream (no GPU visible?)");
      return false;
    }
    if (!upload_grid(hg, g, grid_mem)) return false;
    // device address of the allocation-free entry point's fixed workspace on
    // this device (cfd2lcs_adjoint_forward_noalloc_c reports a failure here)
    ws_ok = scratch_setup();
    if (!status_mem.alloc(sizeof(DevStatus))) {
      fail("hipMalloc failed (status)");
      return false;
    }
    st = (DevStatus*)status_mem.p;
    if (hipHostMalloc((void**)&st_host, sizeof(DevStatus), hipHostMallocDefault) != hipSuccess) {
      fail("hipHostMalloc failed");
      return false;
    }
    std::memset(st_host, 0, sizeof(DevStatus));
    if (hipMemset(st, 0, sizeof(DevStatus)) != hipSuccess) return false;
    for (int b = 0; b < 2; b++) {
      if (!vel[b].alloc(3 * fbytes()) || hipMemset(vel[b].p, 0, 3 * fbytes()) != hipSuccess) {
        fail("hipMalloc failed (velocity)");
        return false;
      }
    }
    if (!stage_u.alloc(3 * sizeof(U) * (size_t)g.np)) {
      fail("hipMalloc failed (staging)");
      return false;
    }
    return true;
  }

  // spline order of the active interpolator, with the order clamp of
  // lp_tracking_m.f90:506-508
  int spline_order_for(int interp) const { return interp == kQuadratic ? 3 : 4; }
  // knots and collocation factors of grid h for the active interpolator
  // While cfd2lcs_adjoint_forward_noalloc_c runs, every buffer the shared
  // code paths would allocate is carved instead from the fixed workspace of
  // scratch.hip (arena != null), and memory is cleared by a fill kernel.
  struct Arena {
    char* base = nullptr;
    size_t size = 0, used = 0;
  };
  Arena* arena = nullptr;
  // b with at least `bytes`: kept if large enough, else allocated (or, in an
  // allocation-free call, carved from the workspace)
  bool get_buf(DevBuf& b, size_t bytes) {
    if (b.p && b.bytes >= bytes) return true;
    if (!arena) return b.alloc(bytes);
    const size_t off = (arena->used + 255) & ~size_t(255);
    if (off + bytes > arena->size) return false;
    b.view(arena->base + off, bytes);
    arena->used = off + bytes;
    return true;
  }
  bool zero_async(void* p, size_t bytes) {
    if (arena) return launch_zero(p, bytes, stream) == hipSuccess;
    return hipMemsetAsync(p, 0, bytes, stream) == hipSuccess;
  }

  bool grid_factors(const HostGrid& h, GridDev<T>& gd, DevBuf sm[3], int& have, int interp) {
    const int k = spline_order_for(interp);
    if (have == k) return true;
    for (int d = 0; d < 3; d++) {
      int kk = std::max(std::min(k, h.N[d] - 1), 2);
      SplineAxis ax;
      std::string e = spline_axis(h.xl[d], kk, ax);
      if (!e.empty()) {
        fail(e);
        return false;
      }
      const size_t nr = ax.knots.size() + ax.band.size() + ax.poly.size();
      std::vector<T> hv(nr);
      for (size_t a = 0; a < ax.knots.size(); a++) hv[a] = (T)ax.knots[a];
      for (size_t a = 0; a < ax.band.size(); a++) hv[ax.knots.size() + a] = (T)ax.band[a];
      for (size_t a = 0; a < ax.poly.size(); a++) hv[ax.knots.size() + ax.band.size() + a] = (T)ax.poly[a];
      if (!get_buf(sm[d], nr * sizeof(T)) ||
          hipMemcpy(sm[d].p, hv.data(), nr * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
        fail("upload of the spline factors failed");
        return false;
      }
      gd.kord[d] = kk;
      gd.kn[d] = (const T*)sm[d].p;
      gd.band[d] = (const T*)sm[d].p + ax.knots.size();
      gd.bpoly[d] = (const T*)sm[d].p + ax.knots.size() + ax.band.size();
    }
    if (!(gd.kord[0] == gd.kord[1] && (gd.kord[2] == gd.kord[0] || gd.kord[2] == 2))) {
      fail("spline orders per axis not supported by this build (need equal x/y orders)");
      return false;
    }
    have = k;
    return true;
  }
  bool ensure_spline(int interp) {
    const int k = spline_order_for(interp);
    if (spline_k == k) return true;
    if (!grid_factors(hg, g, spline_mem, spline_k, interp)) return false;
    vcoef_ok[0] = vcoef_ok[1] = false;
    for (auto& dg : diags)
      for_sets(*dg, [&](Diag& d) { std::fill(d.slot_coef_ok.begin(), d.slot_coef_ok.end(), 0); });
    return true;
  }
  // the spline factors of a set's own grid (same orders as the flow grid's,
  // which the gather kernels are instantiated for)
  bool ensure_set_spline(Diag& dg, int interp) {
    if (!dg.sg) return true;
    SetGrid& sg = *dg.sg;
    if (!grid_factors(sg.hg, sg.g, sg.spline_mem, sg.spline_k, interp)) return false;
    for (int d = 0; d < 3; d++)
      if (sg.g.kord[d] != g.kord[d]) {
        fail("spline order of a diagnostic grid differs from the flow grid's (grid too small along an axis)");
        return false;
      }
    return true;
  }
  bool is_spline(int interp) const { return interp == kQuadratic || interp == kCubic; }

  hipEvent_t new_event() {
    hipEvent_t e = nullptr;
    if (!evpool.empty()) {
      e = evpool.back();
      evpool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
      e = nullptr;
    }
    return e;
  }
  void timed_begin(Timed& t, int cat) {
    t.cat = cat;
    t.a = new_event();
    t.b = new_event();
    if (t.a) (void)hipEventRecord(t.a, stream);
  }
  void timed_end(Timed& t) {
    if (t.b) (void)hipEventRecord(t.b, stream);
    if (t.a && t.b) pending.push_back(t);
  }
  // kernel-level timing, on request only (it adds two events per launch)
  template <class F>
  void ktimed(int cat, F&& launch) {
    if (!g_opt.kernel_timing || arena) {
      launch();
      return;
    }
    Timed t;
    timed_begin(t, cat);
    launch();
    timed_end(t);
  }
  void harvest_timers() {
    for (auto& t : pending) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess && t.cat >= 0 && t.cat < kTimers) {
        tsum[t.cat] += ms * 1e-3;
        tcount[t.cat] += 1;
      }
      evpool.push_back(t.a);
      evpool.push_back(t.b);
    }
    pending.clear();
  }

  // a particle set's own grid from interior coordinates and flags
  // (init_sgrid), uploaded to the device
  std::unique_ptr<SetGrid> make_set_grid(const int nn[3], const std::vector<double> xyz[3],
                                         const std::vector<int>& fl) {
    auto sg = std::make_unique<SetGrid>();
    // generated nodes may leave the coordinate lines by a few ulp
    const double ulp = sizeof(U) == 8 ? DBL_EPSILON : (double)FLT_EPSILON;
    std::string e = sg->hg.build(nn, xyz[0].data(), xyz[1].data(), xyz[2].data(), fl.data(),
                                 sizeof(U) == 8 ? 1e-9 : 1e-6, 64 * ulp);
    if (!e.empty()) {
      fail("diagnostic grid: " + e);
      return nullptr;
    }
    if (sg->hg.n[0] > 32767 || sg->hg.n[1] > 32767 || sg->hg.n[2] > 32767) {
      fail("diagnostic grid too large (more than 32767 nodes along an axis)");
      return nullptr;
    }
    if (!upload_grid(sg->hg, sg->g, sg->mem)) return nullptr;
    return sg;
  }

  // one particle set with its flowmap / slot / composition state
  // (init_lp, lp_m.f90:13-114)
  std::unique_ptr<Diag> new_set(int type, U T_, U h, U dtf, const std::string& label, const std::string& lp_label,
                                std::unique_ptr<SetGrid> sg) {
    auto dg = std::make_unique<Diag>();
    dg->type = type;
    dg->dir = type == 0 ? 1 : -1;
    dg->Tint = T_;
    dg->h = h;
    dg->dt_factor = dtf;
    dg->label = label;
    dg->lp_label = lp_label;
    dg->sg = std::move(sg);
    // nstep = nint(T/h); slots 0..nstep (the write index may equal nstep)
    dg->nstep = (int)std::lround((double)(T_ / h));
    if (dg->nstep < 1 || dg->nstep > 100000) {
      fail("T/h out of range");
      return nullptr;
    }
    dg->nslot = dg->nstep + 1;
    dg->slot_ok.assign(dg->nslot, 0);
    dg->slot_coef_ok.assign(dg->nslot, 0);
    const long long np = NP(*dg);
    const size_t fb = fbytes(NT(*dg)), pb = sizeof(T) * (size_t)np;
    bool ok = true;
    if (dg->dir > 0) {
      ok = ok && dg->part[0].alloc(pset_bytes(np));
    } else {
      // sized for T: the storage may change at any h-interval reset
      ok = ok && dg->fm[0].alloc(3 * fb) && dg->fm[1].alloc(3 * fb);
    }
    ok = ok && dg->slots.alloc((size_t)dg->nslot * 3 * fb) && dg->order.alloc(sizeof(int) * dg->nslot) &&
         dg->dxout.alloc(3 * pb) && dg->compfm.alloc(3 * fb) && dg->comp.alloc(3 * pb) &&
         dg->ftle.alloc(sizeof(double) * (size_t)np);
    if (dg->sg) ok = ok && dg->sg->nodes.alloc(3 * sizeof(int16_t) * (size_t)np);
    if (!ok) {
      fail("hipMalloc failed for a diagnostic (out of device memory?)");
      return nullptr;
    }
    if (hipMemsetAsync(dg->compfm.p, 0, 3 * fb, stream) != hipSuccess ||
        hipMemsetAsync(dg->comp.p, 0, 3 * pb, stream) != hipSuccess ||
        hipMemsetAsync(dg->ftle.p, 0, sizeof(double) * (size_t)np, stream) != hipSuccess) {
      fail("hipMemset failed");
      return nullptr;
    }
    // the tracked flow nodes (without the periodic wrap) of the set's nodes:
    // reset_lp puts the particles back on the same nodes with the same
    // guess, so a backward set's are tracked once; they also locate the
    // velocity at T0 on the set's grid
    if (dg->sg) launch_track_nodes(g, dg->sg->g, (int16_t*)dg->sg->nodes.p, st, stream);
    reset_diag(*dg);
    return dg;
  }

  int add_diag(int type, int resolution, U T_, U h, const char* label) override {
    const std::string lab = label ? label : "";
    if (!hg.rect && (resolution != 0 || g_opt.aux_grid)) {
      fail("resolution != 0 and AUX_GRID on a non-rectilinear grid are not supported by this build");
      return -1;
    }
    // the diagnostic's grid (cfd2lcs.f90:298-310, new_sgrid_from_sgrid)
    std::unique_ptr<SetGrid> sg;
    if (resolution != 0) {
      int nn[3];
      std::vector<double> xyz[3];
      std::vector<int> fl;
      resolution = refine_grid<U>(hg, resolution, nn, xyz, fl);
      if (resolution != 0) {
        sg = make_set_grid(nn, xyz, fl);
        if (!sg) return -1;
      }
    }
    // dt_factor (lcs_m.f90:40-44)
    const U dtf = type == 0 ? (U)1 : (U)1 / (U)std::max(1 + resolution, 1);
    auto dg = new_set(type, T_, h, dtf, lab, lab + "-LP", std::move(sg));
    if (!dg) return -1;
    dg->id = (int)diags.size() + 1;
    if (g_opt.aux_grid) {
      // create_aux_grid (lcs_m.f90:227-435): six sets on shifted copies of
      // the diagnostic's grid, with its flags
      const HostGrid& base = HG(*dg);
      std::vector<int> fl((size_t)base.np);
      for (int k = 1; k <= base.n[2]; k++)
        for (int j = 1; j <= base.n[1]; j++)
          for (int i = 1; i <= base.n[0]; i++) fl[base.pid(i, j, k)] = base.flag[base.id(i, j, k)];
      for (int a = 0; a < 6; a++) {
        std::vector<double> xyz[3];
        aux_grid_coords<U>(base, a, xyz);
        auto ag = make_set_grid(base.n, xyz, fl);
        if (!ag) return -1;
        auto ad = new_set(type, T_, h, dtf, lab, lab + "-LP" + std::to_string(a), std::move(ag));
        if (!ad) return -1;
        dg->aux.push_back(std::move(ad));
      }
    }
    if (hipStreamSynchronize(stream) != hipSuccess) {
      fail("diagnostic initialisation failed on the device");
      return -1;
    }
    if (!sync_status()) return -1;
    diags.push_back(std::move(dg));
    return (int)diags.size();
  }

  // reset_lp (lp_m.f90:228-294) + track_lp2node onto the flow grid:
  // particles back on their nodes, fm = 0
  void reset_diag(Diag& dg) {
    const long long np = NP(dg);
    if (dg.dir > 0) {
      if (dg.sg) launch_fwd_place(g, dg.sg->g, pset(dg.part[dg.pcur], np), st, stream);
      else launch_fwd_reset(g, pset(dg.part[dg.pcur], np), stream);
      dg.sorted = false;
      dg.since_sort = 0;
      dg.cmode = g_opt.compensated;
    } else {
      dg.fm32 = sizeof(T) == 8 && g_opt.fm_storage == kStoreF32;
      (void)hipMemsetAsync(dg.fm[0].p, 0, 3 * dg.fm_elem() * (size_t)NT(dg), stream);
      (void)hipMemsetAsync(dg.fm[1].p, 0, 3 * dg.fm_elem() * (size_t)NT(dg), stream);
      dg.fresh_nodes = true;
    }
  }

  // Time-blended stage velocities u = (1-ct) u_n + ct u_np1 (values, or
  // spline coefficients: the fit is linear) are formed once per distinct ct
  // of the update and shared by all stages and directions that use it.
  struct Blend {
    DevBuf buf;
    U ct = 0;
    int storage = kStoreNative;
    bool spl = false;
    bool valid = false;
    unsigned long long stamp = 0;
  };
  // the pool: six entries, or as many as an update's planned blends need
  // (plan_blends), up to kMaxBlends
  static constexpr int kMaxBlends = 48;
  Blend blends[kMaxBlends];
  int nblend = 6;
  unsigned long long blend_clock = 0;
  int vold = 0;  // index of u_n in vel[] during an update
  double umax[2] = {0.0, 0.0};  // largest |component| of vel[b] (interior)
  // 16-bit stage fields hold u / vscale, vscale = 2^(e-7) for the largest
  // |component| 2^e * m (1 <= m < 2) of both snapshots: stored values stay
  // below 2^8, leaving 2^8 of headroom for spline coefficients under FP16's
  // 65504, and anything above 2^-14 * vscale stays a normal number
  double vscale = 1.0;

  struct StageField {
    const void* c[3];
  };
  // stage velocity at blend weight ct, in the storage the gathers read
  StageField stage_field(U ct, bool spl) {
    DevBuf& A = spl ? vcoef[vold] : vel[vold];
    DevBuf& B = spl ? vcoef[cur] : vel[cur];
    const int stor = g_opt.vel_storage;
    // the two snapshots' fields as the blend reads them
    const bool in32 = spl && vcoef32[vold];
    auto whole = [&](DevBuf& b) {
      if (in32) {
        const CF3<float> f = cf3e<float>(b, g.nt);
        return StageField{{f.c[0], f.c[1], f.c[2]}};
      }
      CF3<T> f = cf3(b);
      return StageField{{f.c[0], f.c[1], f.c[2]}};
    };
    if (stor == kStoreNative || (stor == kStoreF32 && in32)) {
      if (ct == (U)0) return whole(A);
      if (ct == (U)1) return whole(B);
    }
    // per-component stride of the stage field (0 for the packed records)
    const size_t eb = stor == kStoreNative ? sizeof(T) : (stor == kStoreF32 ? sizeof(float) : 0);
    const size_t fbytes_stage = stage_field_bytes(stor, sizeof(T), (size_t)g.nt);
    // A rounded storage format reuses a blend whose weight is within 2^-10
    // of the format's unit roundoff of ct: the forward and backward
    // sub-cycles reach the same stage times by sums rounded differently, and
    // the velocity such a weight difference makes is far below the rounding
    // of the stored value.  Native storage matches ct exactly.
    const double tol = stor == kStoreF32 ? 0x1p-34 : (stor == kStoreF16 ? 0x1p-21 : (stor == kStoreBF16 ? 0x1p-18 : 0.0));
    Blend* victim = nullptr;
    for (int q = 0; q < nblend; q++) {
      Blend& b = blends[q];
      if (b.valid && std::fabs((double)b.ct - (double)ct) <= tol && b.storage == stor && b.spl == spl) {
        b.stamp = ++blend_clock;
        char* p = (char*)b.buf.p;
        return StageField{{p, p + eb * g.nt, p + 2 * eb * g.nt}};
      }
      // an allocation-free call reuses only the entries it gave workspace to
      if (arena && !b.buf.p) continue;
      if (!victim || (!b.valid && victim->valid) || (b.valid == victim->valid && b.stamp < victim->stamp)) victim = &b;
    }
    if (!victim || (!(victim->buf.p && victim->buf.bytes >= fbytes_stage) && !get_buf(victim->buf, 3 * fbytes()))) {
      fail("hipMalloc failed (blended velocity)");
      return whole(A);
    }
    char* p = (char*)victim->buf.p;
    void* out[3] = {p, p + eb * g.nt, p + 2 * eb * g.nt};
    if constexpr (sizeof(T) == 8) {
      if (in32) {
        ktimed(14, [&] {
          launch_blend(g, out, stor, cf3e<float>(A, g.nt), cf3e<float>(B, g.nt), (T)ct, (T)(1.0 / vscale), st, stream);
        });
      } else {
        ktimed(14, [&] { launch_blend(g, out, stor, cf3(A), cf3(B), (T)ct, (T)(1.0 / vscale), st, stream); });
      }
    } else {
      ktimed(14, [&] { launch_blend(g, out, stor, cf3(A), cf3(B), (T)ct, (T)(1.0 / vscale), st, stream); });
    }
    victim->ct = ct;
    victim->storage = stor;
    victim->spl = spl;
    victim->valid = true;
    victim->stamp = ++blend_clock;
    return StageField{{out[0], out[1], out[2]}};
  }

  // ct of a stage at time tt (lp_motion_m.f90:170 etc.); NaN (first update)
  // clamps to 1 as with IEEE minNum/maxNum
  static U ctf(U tt, U t0, U t1) { return std::fmax(std::fmin((tt - t0) / (t1 - t0), (U)1), (U)0); }

  // the blend weights ct of the stages of a sub-cycle from t by dt (c[0] at
  // the start; RK4's c[2] is its c[1]); returns the stage count
  int stage_cts(U t, U dt, U c[4]) const {
    const U half = (U)1 / (U)2;
    c[0] = c[1] = c[2] = c[3] = 0;
    int ns = 1;
    switch (g_opt.integrator) {
      case kEuler: c[0] = ctf(t, t_n, t_np1); break;
      case kTrapezoidal: c[0] = ctf(t + half * dt, t_n, t_np1); break;
      case kRK2:
        c[0] = ctf(t, t_n, t_np1);
        c[1] = ctf(t + half * dt, t_n, t_np1);
        ns = 2;
        break;
      case kRK3:
        c[0] = ctf(t, t_n, t_np1);
        c[1] = ctf(t + half * dt, t_n, t_np1);
        c[2] = ctf(t + dt, t_n, t_np1);
        ns = 3;
        break;
      default:
        c[0] = ctf(t, t_n, t_np1);
        c[1] = ctf(t + half * dt, t_n, t_np1);
        c[2] = c[1];
        c[3] = ctf(t + dt, t_n, t_np1);
        ns = 4;
    }
    return ns;
  }

  StageParams<T> stage_params(U t, U dt, bool spl) {
    StageParams<T> sp{};
    sp.integ = g_opt.integrator;
    sp.dt = (T)dt;
    U c[4];
    const int ns = stage_cts(t, dt, c);
    sp.storage = g_opt.vel_storage;
    sp.vscale = (T)vscale;
    sp.lds = g_opt.spline_lds;
    for (int a = 0; a < 4; a++) {
      const StageField f = stage_field(c[a < ns ? a : 0], spl);
      for (int d = 0; d < 3; d++) sp.v[a][d] = f.c[d];
    }
    sp.un = cf3(vel[vold]);
    sp.unp1 = cf3(vel[cur]);
    sp.ct0 = (T)c[0];
    return sp;
  }

  void update(const U* u, const U* v, const U* w, bool device, U time) override {
    int interp = g_opt.interpolator;
    if (sizeof(T) == 8 && (g_opt.vel_storage == kStoreF16 || g_opt.vel_storage == kStoreBF16)) {
      fail("VELOCITY_STORAGE = LCS_FP16 / LCS_BF16 needs PRECISION = LCS_FP32");
      return;
    }
    if (interp != kNearest && interp != kLinear && interp != kQuadratic && interp != kCubic && interp != kTSE &&
        interp != kTSELimit) {
      fail("Unrecognized interpolator");
      return;
    }
    if (g_opt.integrator < kEuler || g_opt.integrator > kRK4) {
      fail("unknown INTEGRATOR");
      return;
    }
    // a non-rectilinear grid interpolates with TSE, or TSE_LIMIT for every
    // other choice (lp_tracking_m.f90:319-328)
    if (!hg.rect && interp != kTSE) interp = kTSELimit;
    const bool spl = is_spline(interp);
    if (spl && !ensure_spline(interp)) return;
    // time levels (cfd2lcs.f90:104-111) and u_n <- u_np1 (112)
    if (first_call) {
      t_n = time;
      t_np1 = time;
      first_call = false;
    } else {
      t_n = t_np1;
      t_np1 = time;
    }
    const int old = cur;
    cur = 1 - cur;
    vold = old;
    vcoef_ok[cur] = false;
    for (auto& b : blends) b.valid = false;
    // vel[cur] becomes u_np1; vel[old] (previous u_np1) is u_n; vcoef[b]
    // holds the spline coefficients of vel[b]
    const U *su = u, *sv = v, *sw = w;
    if (!device) {
      U* s = (U*)stage_u.p;
      const size_t b = sizeof(U) * (size_t)g.np;
      HIPCHK(hipMemcpyAsync(s, u, b, hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(s + g.np, v, b, hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(s + 2 * g.np, w, b, hipMemcpyHostToDevice, stream));
      su = s;
      sv = s + g.np;
      sw = s + 2 * g.np;
    }
    F3<T> unp1 = f3(vel[cur]);
    CF3<T> unm = cf3(vel[old]);
    HIPCHK(hipMemsetAsync(&st->cfl_bits, 0, sizeof(st->cfl_bits) + sizeof(st->umax_bits), stream));
    ktimed(13, [&] {
      launch_ingest<U, T>(g, unp1, su, sv, sw, stream);
      launch_vel_ghost(g, unp1, unm, st, stream);
      launch_cfl(g, unp1, st, stream);
    });
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(st_host, st, sizeof(DevStatus), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    harvest_timers();
    oob_total = st_host->oob;
    if (st_host->err) {
      fail(st_host->err & kErrBcFlag
               ? "unknown bcflag"
               : (st_host->err & kErrTrack
                      ? "particle tracking failed"
                      : (st_host->err & kErrRange ? "velocity outside the FP16 range of VELOCITY_STORAGE = LCS_FP16"
                                                  : "device error")));
      return;
    }
    double cfl_rate = 0;
    {
      unsigned long long bits = st_host->cfl_bits;
      std::memcpy(&cfl_rate, &bits, sizeof(double));
      bits = st_host->umax_bits;
      std::memcpy(&umax[cur], &bits, sizeof(double));
      const double m = std::max(umax[0], umax[1]);
      vscale = 1.0;
      if (g_opt.vel_storage == kStoreF16 && m > 0.0 && std::isfinite(m)) vscale = std::ldexp(1.0, std::ilogb(m) - 7);
    }
    if (spl) {
      // spline coefficients of u_np1 (and of u_n when they are missing or in
      // the other element type)
      if (!fit_vel(cur)) return;
      if ((!vcoef_ok[old] || vcoef32[old] != want_vcoef32()) && !fit_vel(old)) return;
    }
    const U dt_f = t_np1 - t_n;
    plan_blends(spl, cfl_rate, dt_f);
    if (g_error) return;
    for (auto& dgp : diags) {
      for_sets(*dgp, [&](Diag& dg) {
        if (!g_error) advect(dg, interp, spl, cfl_rate, dt_f);
      });
      if (g_error) return;
    }
    // diagnostics at h boundaries (cfd2lcs.f90:144-224)
    for (auto& dgp : diags) {
      Diag& dg = *dgp;
      if ((long long)(t_np1 / dg.h) == (long long)(t_n / dg.h)) continue;
      output(dg, interp);
      if (g_error) return;
    }
  }

  // one particle set through the CFD step: set_dt (lp_motion_m.f90:56-127),
  // then update_lp (forward) or update_flowmap_sl (backward)
  // set_dt (lp_motion_m.f90:90-101): the sub-cycle count of a set through
  // the CFD step (0 when above 1e6) and its signed sub-cycle step
  int sub_cycle_count(const Diag& dg, double cfl_rate, U dt_f, U& dt) const {
    const U cmax = (U)((T)cfl_rate) * dt_f;
    const double r = std::ceil((double)(cmax / (g_opt.cfl_max * dg.dt_factor)));
    if (r > 1e6) return 0;
    const int nsub = r >= 1 ? (int)r : 1;
    dt = dt_f / (U)nsub;
    if (dg.dir < 0) dt = (U)-1 * dt;
    return nsub;
  }

  // Every stage blend the update's sub-cycles will ask for, formed in
  // batches that read the two fields once per kBlendBatch blends, before the
  // sweeps: the weights come from stage_cts() along the same sub-cycle times
  // the sweeps step through, so the sweeps find them in the pool by the
  // pool's own match (stage_field).  A weight this misses is blended on
  // demand.
  void plan_blends(bool spl, double cfl_rate, U dt_f) {
    nblend = 6;
    const int stor = g_opt.vel_storage;
    if (arena) return;
    const bool in32 = spl && vcoef32[vold];
    const bool snap01 = stor == kStoreNative || (stor == kStoreF32 && in32);
    const double tol = stor == kStoreF32 ? 0x1p-34 : (stor == kStoreF16 ? 0x1p-21 : (stor == kStoreBF16 ? 0x1p-18 : 0.0));
    std::vector<U> cts;
    for (auto& dgp : diags) {
      U dt = 0;
      const int nsub = sub_cycle_count(*dgp, cfl_rate, dt_f, dt);
      U t = dgp->dir > 0 ? t_n : t_np1;
      for (int s = 0; s < nsub; s++) {
        U c[4];
        const int ns = stage_cts(t, dt, c);
        for (int a = 0; a < 4; a++) {
          const U ct = c[a < ns ? a : 0];
          if (snap01 && (ct == (U)0 || ct == (U)1)) continue;
          bool have = false;
          for (U x : cts) have = have || std::fabs((double)x - (double)ct) <= tol;
          if (!have) cts.push_back(ct);
        }
        t = t + dt;
      }
    }
    // at most kMaxBlends - 2 planned entries (two stay for on-demand
    // blends) and at most 16 GiB of them
    const size_t eb = 3 * fbytes();
    const size_t fit = std::min<size_t>(kMaxBlends - 2, ((size_t)16 << 30) / std::max<size_t>(eb, 1));
    if (cts.size() < 2 || cts.size() > fit) return;
    nblend = std::max(6, (int)cts.size() + 2);
    // per-component stride of the stage field (0: one array of packed records)
    const size_t sb = stor == kStoreNative ? sizeof(T) : (stor == kStoreF32 ? sizeof(float) : 0);
    DevBuf& A = spl ? vcoef[vold] : vel[vold];
    DevBuf& B = spl ? vcoef[cur] : vel[cur];
    for (size_t q0 = 0; q0 < cts.size(); q0 += kBlendBatch) {
      BlendBatch<T> bb{};
      bb.inv = (T)(1.0 / vscale);
      bb.n = (int)std::min<size_t>(kBlendBatch, cts.size() - q0);
      for (int q = 0; q < bb.n; q++) {
        Blend& e = blends[q0 + q];
        if (!(e.buf.p && e.buf.bytes >= eb) && !get_buf(e.buf, eb)) {
          fail("hipMalloc failed (blended velocity)");
          return;
        }
        char* p = (char*)e.buf.p;
        for (int d = 0; d < 3; d++) bb.out[q][d] = p + d * sb * (size_t)g.nt;
        bb.ct[q] = (T)cts[q0 + q];
        e.ct = cts[q0 + q];
        e.storage = stor;
        e.spl = spl;
        e.valid = true;
        e.stamp = ++blend_clock;
      }
      if constexpr (sizeof(T) == 8) {
        if (in32) {
          ktimed(14, [&] { launch_blend_batch(g, bb, stor, cf3e<float>(A, g.nt), cf3e<float>(B, g.nt), st, stream); });
          continue;
        }
      }
      ktimed(14, [&] { launch_blend_batch(g, bb, stor, cf3(A), cf3(B), st, stream); });
    }
  }

  void advect(Diag& dg, int interp, bool spl, double cfl_rate, U dt_f) {
    U dt = 0;
    const int nsub = sub_cycle_count(dg, cfl_rate, dt_f, dt);
    if (nsub == 0) {
      fail("CFL sub-cycle count above 1e6 (velocity or time step out of range)");
      return;
    }
    nsub_last = nsub;
    const long long np = NP(dg), nt = NT(dg);
    if (std::fabs(dt) > 0) {
      if (dg.dir > 0) integ_fwd += np * nsub;
      else integ_bkwd += np * nsub;
    }
    Timed tm;
    timed_begin(tm, dg.dir > 0 ? 0 : 1);
    if (dg.dir > 0) {
      U t = t_n;
      if (g_opt.particle_sort > 0 && dg.since_sort >= g_opt.particle_sort) {
        bool ok = true;
        ktimed(15, [&] { ok = sort_particles(dg); });
        if (!ok) return;
      }
      const ParticleSet<T> ps = pset(dg.part[dg.pcur], np);
      for (int s = 0; s < nsub; s++) {
        const StageParams<T> sp = stage_params(t, dt, spl);
        ktimed(8, [&] { launch_fwd_step(g, interp, sp, ps, (int)np, dg.cmode, st, stream); });
        t = t + dt;
      }
      if (std::fabs(dt) > 0) dg.since_sort++;
    } else {
      U t = t_np1;
      GridDev<T>& lg = G(dg);
      if (spl && !ensure_set_spline(dg, interp)) return;
      if (spl && !dg.fmcoef.p && !dg.fmcoef.alloc(3 * fbytes(nt))) {
        fail("hipMalloc failed (flowmap coefficients)");
        return;
      }
      // the MASK test reads lp%no_scfd as it stands at the start of the update
      const bool fresh_at_start = dg.fresh_nodes;
      auto sub_cycles = [&](auto tag) {
        using E = decltype(tag);
        for (int s = 0; s < nsub; s++) {
          DevBuf& a = dg.fm[dg.fcur];
          DevBuf& b = dg.fm[1 - dg.fcur];
          CF3<E> src = cf3e<E>(a, nt);
          if (spl) {
            // coefficients of fm into fmcoef (the copy rides on the first pass)
            ktimed(10, [&] { launch_spline_fit(lg, f3e<E>(dg.fmcoef, nt), 3, stream, src.c); });
            src = cf3e<E>(dg.fmcoef, nt);
          }
          const StageParams<T> sp = stage_params(t, dt, spl);
          if (dg.sg) {
            const int fresh = (dg.fresh_nodes ? 1 : 0) | (fresh_at_start ? 2 : 0);
            ktimed(9, [&] {
              launch_sl_step2(g, lg, (const int16_t*)dg.sg->nodes.p, fresh, interp, sp, src, f3e<E>(b, nt), st,
                              stream);
            });
            dg.fresh_nodes = false;
          } else {
            ktimed(9, [&] { launch_sl_step(g, interp, sp, src, f3e<E>(b, nt), st, stream, g_opt.sl_tiled); });
          }
          ktimed(13, [&] { launch_fm_ghost(lg, f3e<E>(b, nt), cf3e<E>(a, nt), st, stream); });
          dg.fcur = 1 - dg.fcur;
          t = t + dt;
        }
      };
      if constexpr (sizeof(T) == 8) {
        if (dg.fm32) sub_cycles(float{});
        else sub_cycles(T{});
      } else {
        sub_cycles(T{});
      }
    }
    timed_end(tm);
    HIPCHK(hipGetLastError());
  }

  bool fit_vel(int b) {
    if (!get_buf(vcoef[b], 3 * fbytes())) {
      fail("hipMalloc failed (velocity coefficients)");
      return false;
    }
    const CF3<T> u = cf3(vel[b]);
    const bool w32 = want_vcoef32();
    if constexpr (sizeof(T) == 8) {
      if (w32) ktimed(10, [&] { launch_spline_fit(g, f3e<float>(vcoef[b], g.nt), 3, stream, u.c); });
      else ktimed(10, [&] { launch_spline_fit(g, f3(vcoef[b]), 3, stream, u.c); });
    } else {
      ktimed(10, [&] { launch_spline_fit(g, f3(vcoef[b]), 3, stream, u.c); });
    }
    vcoef_ok[b] = true;
    vcoef32[b] = w32;
    return true;
  }

  T* slot_ptr(DevBuf& b, int s, long long nt) { return (T*)b.p + (size_t)s * 3 * nt; }

  // write_flowmap_substep (flowmap_m.f90:106-156) of one set: the h-interval
  // flowmap into ring slot `step` (forward sets: exchange_lpmap first)
  void store_slot(Diag& d, int step) {
    const long long nt = NT(d);
    T* sp = slot_ptr(d.slots, step, nt);
    F3<T> slot{{sp, sp + nt, sp + 2 * nt}};
    if (d.dir > 0) {
      launch_field_from_origin(G(d), fwd_dx_by_origin(d), slot, stream);
    } else if (d.fm32) {
      launch_widen((const float*)d.fm[d.fcur].p, sp, 3 * nt, stream);
    } else {
      HIPCHK(hipMemcpyAsync(sp, d.fm[d.fcur].p, 3 * fbytes(nt), hipMemcpyDeviceToDevice, stream));
    }
    d.slot_ok[step] = 1;
    d.slot_coef_ok[step] = 0;
    if (g_opt.write_slots) (void)write_slot_files(d, step);
  }

  // reconstruct_flowmap (flowmap_m.f90:158-323) of one set on its grid:
  // composed displacement by origin node into dxout, as a ghosted field in
  // compfm
  void compose_set(Diag& d, int interp, bool spl, int step) {
    const long long nt = NT(d), np = NP(d);
    GridDev<T>& lg = G(d);
    if (spl) {
      if (!ensure_set_spline(d, interp)) return;
      if (!d.slotcoef.p && !d.slotcoef.alloc((size_t)d.nslot * 3 * fbytes(nt))) {
        fail("hipMalloc failed (slot coefficients)");
        return;
      }
      for (int s = 0; s < d.nslot; s++) {
        if (!d.slot_ok[s] || d.slot_coef_ok[s]) continue;
        T* cp = slot_ptr(d.slotcoef, s, nt);
        const T* sp = slot_ptr(d.slots, s, nt);
        const T* in[3] = {sp, sp + nt, sp + 2 * nt};
        ktimed(10, [&] { launch_spline_fit(lg, F3<T>{{cp, cp + nt, cp + 2 * nt}}, 3, stream, in); });
        d.slot_coef_ok[s] = 1;
      }
    }
    // slot order (flowmap_m.f90:202-268)
    std::vector<int>& order = d.order_h;
    order.clear();
    int s = step;
    const int inc = d.dir > 0 ? 1 : -1;
    if (d.dir > 0) s = s + inc;
    for (int it = 0; it < d.nstep; it++) {
      if (s >= d.nstep) s = 0;
      if (s < 0) s = d.nstep - 1;
      if (d.slot_ok[s]) order.push_back(s);
      s += inc;
    }
    if (g_opt.compose_sort > 0 && !order.empty()) {
      if (!compose_sorted(d, interp, (const T*)(spl ? d.slotcoef.p : d.slots.p), order)) return;
    } else {
      HIPCHK(hipMemcpyAsync(d.order.p, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice, stream));
      ktimed(11, [&] {
        launch_compose(lg, interp, (const T*)(spl ? d.slotcoef.p : d.slots.p), nt, (const int*)d.order.p,
                       (int)order.size(), p3(d.dxout, np), st, stream);
      });
    }
    ktimed(13, [&] { launch_field_from_origin(lg, cp3(d.dxout, np), f3(d.compfm, nt), stream); });
  }

  // the composition one slot per launch (k_compose_slot), the particles
  // re-ordered by the brick of their tracked node before every
  // COMPOSE_SORT-th slot; same results as k_compose
  bool compose_sorted(Diag& d, int interp, const T* base, const std::vector<int>& order) {
    GridDev<T>& lg = G(d);
    const long long np = NP(d), nt = NT(d);
    const int ns = (int)order.size();
    const int every = g_opt.compose_sort;
    // the brick keys of k_compose_slot (comp_key): 8 x 8 x 4 nodes
    const long long nbricks = (long long)((lg.n[0] + 7) / 8) * ((lg.n[1] + 7) / 8) * ((lg.n[2] + 3) / 4);
    const size_t rb = sizeof(CompRec<T>) * (size_t)np;
    bool ok = ns < 2 || (crec[0].ensure(rb) && crec[1].ensure(rb));
    const bool sorts = ns > 1 && every <= ns - 1;
    size_t tb = 0;
    if (ok && sorts) {
      tb = compose_sort_temp_bytes(np, nbricks);
      ok = tb && ckeys.ensure(2 * sizeof(unsigned) * (size_t)np) && cvals.ensure(2 * sizeof(unsigned) * (size_t)np) &&
           ctemp.ensure(tb);
    }
    if (!ok) {
      fail("hipMalloc failed (composition records)");
      return false;
    }
    int cur = 0;
    const unsigned* perm = nullptr;
    for (int s = 0; s < ns; s++) {
      const T* sb = base + (size_t)order[s] * 3 * (size_t)nt;
      const CF3<T> f{{sb, sb + nt, sb + 2 * nt}};
      const bool last = s == ns - 1;
      const bool sort_next = !last && (s + 1) % every == 0;
      unsigned* keys = sort_next ? (unsigned*)ckeys.p : nullptr;
      unsigned* vals = sort_next ? (unsigned*)cvals.p : nullptr;
      ktimed(11, [&] {
        launch_compose_slot(lg, interp, f, (const CompRec<T>*)crec[cur].p, perm, (CompRec<T>*)crec[1 - cur].p, keys,
                            vals, s == 0 ? 1 : 0, last ? 1 : 0, p3(d.dxout, np), st, stream);
      });
      if (last) break;
      perm = nullptr;
      if (sort_next) {
        ktimed(16, [&] { perm = launch_compose_sort(keys, vals, np, nbricks, ctemp.p, tb, stream); });
        if (!perm) {
          fail("composition sort failed");
          return false;
        }
      }
      cur = 1 - cur;
    }
    HIPCHK_R(hipGetLastError(), false);
    return true;
  }

  // the diagnostic at an h boundary (cfd2lcs.f90:147-213): every set's
  // substep and composition, compute_ftle, the output file, reset_lp
  void output(Diag& dg, int interp) {
    const double md = std::fmod((double)t_np1, (double)dg.Tint);
    const int step = (int)std::lround((double)((U)md / dg.h));
    if (step < 0 || step >= dg.nslot) {
      fail("flowmap slot index out of range");
      return;
    }
    for_sets(dg, [&](Diag& d) {
      if (!g_error) store_slot(d, step);
    });
    if (g_error || !io_ok()) return;
    if (g_opt.write_lcs && (g_opt.velocity_io || g_opt.invariants) && any_fwd() && !record_velocity(step)) return;
    const bool spl = is_spline(interp);
    Timed tc;
    timed_begin(tc, 2);
    for_sets(dg, [&](Diag& d) {
      if (!g_error) compose_set(d, interp, spl, step);
    });
    timed_end(tc);
    if (g_error) return;
    Timed tf;
    timed_begin(tf, 3);
    const double tiny = sizeof(U) == 8 ? DBL_MIN : (double)FLT_MIN;
    GridDev<T>& lg = G(dg);
    const long long np = NP(dg), nt = NT(dg);
    if (dg.aux.empty()) {
      ktimed(12, [&] {
        launch_ftle(lg, cf3(dg.compfm, nt), (double)dg.Tint, g_opt.incompressible, tiny, (double*)dg.ftle.p,
                    p3(dg.comp, np), stream);
      });
    } else {
      // gradient from the six auxiliary sets (lcs_m.f90:95-126); the FM
      // columns are the diagnostic's own composed flowmap
      AuxSets<T> ax;
      for (int a = 0; a < 6; a++) {
        Diag& A = *dg.aux[a];
        for (int c = 0; c < 3; c++) ax.disp[a][c] = (const T*)A.dxout.p + (size_t)c * np;
        ax.line[a] = A.sg->g.gl[a % 3];
      }
      ktimed(12, [&] {
        launch_ftle_aux(lg, ax, 100.0 * tiny, (double)dg.Tint, g_opt.incompressible, tiny, (double*)dg.ftle.p,
                        stream);
      });
      HIPCHK(hipMemcpyAsync(dg.comp.p, dg.dxout.p, 3 * sizeof(T) * (size_t)np, hipMemcpyDeviceToDevice, stream));
    }
    timed_end(tf);
    for_sets(dg, [&](Diag& d) { reset_diag(d); });
    dg.noutput++;
    HIPCHK(hipGetLastError());
    // the order lists must stay valid until their copies ran: wait here
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipMemcpy(st_host, st, sizeof(DevStatus), hipMemcpyDeviceToHost));
    if (st_host->err & kErrTrack) {
      fail("cannot locate all particles after the maximum number of tracking iterations");
      return;
    } else if (st_host->err) {
      fail("device error during composition");
      return;
    }
    if (g_opt.write_lcs) write_lcs_file(dg, step);
  }

  // ---- output files ----------------------------------------------------
  io::Writer& wr() {
    if (!writer) writer = std::make_unique<io::Writer>(2);
    return *writer;
  }
  // a failure of an earlier asynchronous write sets the error flag here
  bool io_ok() {
    std::string m;
    if (writer && writer->failed(&m)) {
      fail(m);
      return false;
    }
    return true;
  }
  int flush_output() override {
    if (writer) writer->flush();
    return io_ok() ? 0 : -1;
  }
  bool any_fwd() const {
    for (auto& d : diags)
      if (d->dir > 0) return true;
    return false;
  }
  // n device values of T (on the stream) as a host array of lcsdata_t
  std::shared_ptr<std::vector<U>> host_copy(const T* d, size_t n) {
    auto h = std::make_shared<std::vector<U>>(n);
    if (std::is_same<T, U>::value) {
      if (hipMemcpyAsync(h->data(), d, n * sizeof(T), hipMemcpyDeviceToHost, stream) != hipSuccess ||
          hipStreamSynchronize(stream) != hipSuccess)
        return nullptr;
    } else {
      std::vector<T> t(n);
      if (hipMemcpyAsync(t.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, stream) != hipSuccess ||
          hipStreamSynchronize(stream) != hipSuccess)
        return nullptr;
      for (size_t a = 0; a < n; a++) (*h)[a] = (U)t[a];
    }
    return h;
  }
  // interior of a ghosted 3-component field, on the host
  std::shared_ptr<std::vector<U>> host_interior(CF3<T> src) { return host_interior(g, src); }
  std::shared_ptr<std::vector<U>> host_interior(const GridDev<T>& gd, CF3<T> src) {
    if (!iscratch.ensure(3 * sizeof(T) * (size_t)gd.np)) return nullptr;
    launch_interior(gd, src, p3(iscratch, gd.np), stream);
    return host_copy((const T*)iscratch.p, 3 * (size_t)gd.np);
  }
  static void add3(io::Job& j, const std::shared_ptr<std::vector<U>>& h, size_t np) {
    for (int d = 0; d < 3; d++) j.arrays.push_back(io::Array{h->data() + d * np, np, sizeof(U) == 8});
    j.keep.push_back(h);
  }
  // the reference's file names are character(len=128) variables
  static std::string fname128(const std::string& s) { return s.size() > 128 ? s.substr(0, 128) : s; }

  // write_flowmap_substep (flowmap_m.f90:106-156): the h-interval flowmap of
  // slot `step` (on the set's grid, file <lp%label>-FM) and, with
  // VELOCITY_IO, u_np1 on the flow grid; every set of a diagnostic writes
  // the same u_np1 file, so the diagnostic's own set writes it once
  bool write_slot_files(Diag& dg, int step) {
    std::string e;
    if (!io::ensure_dir("cfd2lcs_tmp", e)) {
      fail(e);
      return false;
    }
    const long long nt = NT(dg);
    const size_t np = (size_t)NP(dg);
    T* sp = slot_ptr(dg.slots, step, nt);
    auto h = host_interior(G(dg), CF3<T>{{sp, sp + nt, sp + 2 * nt}});
    if (!h) {
      fail("copy of a flowmap slot to the host failed");
      return false;
    }
    io::Job j;
    j.kind = io::Job::kUnformatted;
    j.path = fname128("./cfd2lcs_tmp/" + io::i4(0) + "_" + dg.lp_label + "-FM_" + io::i4(step) + ".dat");
    add3(j, h, np);
    wr().submit(std::move(j));
    if ((g_opt.velocity_io || g_opt.invariants) && dg.lp_label == dg.label + "-LP") {
      const size_t np = (size_t)g.np;
      auto hv = host_interior(cf3(vel[cur]));
      if (!hv) {
        fail("copy of the velocity to the host failed");
        return false;
      }
      io::Job jv;
      jv.kind = io::Job::kUnformatted;
      jv.path = fname128("./cfd2lcs_tmp/" + io::i4(0) + "_U_NP1_" + io::i4(step) + ".dat");
      add3(jv, hv, np);
      wr().submit(std::move(jv));
    }
    return true;
  }

  bool record_velocity(int step) {
    if ((int)uring.size() <= step) {
      uring.resize((size_t)step + 1);
      uring_ok.resize((size_t)step + 1, 0);
    }
    if (!uring[step].p && !uring[step].alloc(3 * sizeof(T) * (size_t)g.np)) {
      fail("hipMalloc failed (velocity ring for the FTLE output)");
      return false;
    }
    launch_interior(g, cf3(vel[cur]), p3(uring[step]), stream);
    uring_ok[step] = 1;
    return true;
  }

  // write_lcs + write_ftle (io_nohdf5_m.f90:47-265): X, Y, Z, [FM-X..Z],
  // [U-X..Z], [FLAG], FTLE of the latest composition
  // lp%ugrid of a set on its own grid (flowmap_m.f90:228-247): the interior
  // flow-grid velocity du, read back with periodic ghosts and zero
  // elsewhere, interpolated at the set's nodes through their tracked
  // (clamped) flow nodes with the active interpolator
  const T* velocity_on_set(Diag& dg, const T* du) {
    const int interp = g_opt.interpolator;
    if (!uscratch[0].ensure(3 * fbytes()) || !uscratch[1].ensure(3 * sizeof(T) * (size_t)NP(dg))) {
      fail("hipMalloc failed (velocity on a diagnostic grid)");
      return nullptr;
    }
    launch_expand_periodic(g, du, f3(uscratch[0]), stream);
    if (is_spline(interp)) {
      if (!ensure_spline(interp)) return nullptr;
      launch_spline_fit(g, f3(uscratch[0]), 3, stream);
    }
    launch_field_at_nodes(g, G(dg), (const int16_t*)dg.sg->nodes.p, interp, cf3(uscratch[0]), (T*)uscratch[1].p,
                          stream);
    return (const T*)uscratch[1].p;
  }

  void write_lcs_file(Diag& dg, int step) {
    const HostGrid& hgd = HG(dg);
    const GridDev<T>& lg = G(dg);
    const size_t np = (size_t)lg.np;
    long long findex = std::llround((double)(U)(t_np1 / dg.h));
    if (dg.dir > 0) findex -= std::llround((double)(U)(dg.Tint / dg.h));
    std::string fs;
    if (!io::findex_string(findex, fs)) {
      fail("unsupported range for file index, " + std::to_string(findex));
      return;
    }
    std::string e;
    if (!io::ensure_dir("cfd2lcs_output", e)) {
      fail(e);
      return;
    }
    std::shared_ptr<std::vector<U>>& oxyz = dg.sg ? dg.sg->out_xyz : out_xyz;
    std::shared_ptr<std::vector<U>>& oflag = dg.sg ? dg.sg->out_flag : out_flag;
    if (!oxyz) {
      auto v = std::make_shared<std::vector<U>>(3 * np);
      auto f = std::make_shared<std::vector<U>>(np);
      for (int k = 1; k <= hgd.n[2]; k++)
        for (int j = 1; j <= hgd.n[1]; j++)
          for (int i = 1; i <= hgd.n[0]; i++) {
            const long long p = hgd.pid(i, j, k), l = hgd.id(i, j, k);
            for (int d = 0; d < 3; d++) (*v)[d * np + p] = (U)hgd.g[d][l];
            (*f)[p] = (U)hgd.flag[l];
          }
      oxyz = v;
      oflag = f;
    }
    io::Job j;
    j.kind = io::Job::kTecplot;
    j.path = fname128("./cfd2lcs_output/" + dg.label + "_" + fs + ".dat");
    const std::string ft = fname128(dg.label + "-FTLE");
    std::string hd = "TITLE = \"" + ft + "\"\nVARIABLES =\n\"X\"\n\"Y\"\n\"Z\"\n";
    add3(j, oxyz, np);
    if (g_opt.flowmap_io) {
      hd += "\"FM-X\"\n\"FM-Y\"\n\"FM-Z\"\n";
      auto h = host_copy((const T*)dg.comp.p, 3 * np);
      if (!h) {
        fail("copy of the composed flowmap to the host failed");
        return;
      }
      add3(j, h, np);
    }
    // lp%ugrid after reconstruct_flowmap (flowmap_m.f90:212-247): the
    // velocity read back from slot `step` (BKWD, just written: u_np1) or
    // step + 1 (FWD: the oldest snapshot of the ring); zero when that slot
    // was never written
    const T* du = nullptr;
    if (g_opt.velocity_io || g_opt.invariants) {
      if (dg.dir < 0) {
        if (!vint.p && !vint.alloc(3 * sizeof(T) * (size_t)g.np)) {
          fail("hipMalloc failed (velocity for the FTLE output)");
          return;
        }
        launch_interior(g, cf3(vel[cur]), p3(vint), stream);
        du = (const T*)vint.p;
      } else {
        const int vs = step + 1;
        if (vs < (int)uring.size() && uring_ok[vs]) du = (const T*)uring[vs].p;
      }
      // on a diagnostic grid: interpolated at its nodes (the reference reads
      // the flow-grid file with the diagnostic grid's extents here)
      if (du && dg.sg && !(du = velocity_on_set(dg, du))) return;
    }
    if (g_opt.velocity_io) {
      hd += "\"U-X\"\n\"U-Y\"\n\"U-Z\"\n";
      auto h = du ? host_copy(du, 3 * np) : std::make_shared<std::vector<U>>(3 * np, (U)0);
      if (!h) {
        fail("copy of the velocity to the host failed");
        return;
      }
      add3(j, h, np);
    }
    if (g_opt.bcflag_io) {
      hd += "\"FLAG\"\n";
      j.arrays.push_back(io::Array{oflag->data(), np, sizeof(U) == 8});
      j.keep.push_back(oflag);
    }
    if (g_opt.invariants && !lg.rect) {
      fail("VELOCITY_INVARIANTS on a non-rectilinear grid (least-squares velocity gradient) is not supported "
           "by this build");
      return;
    }
    if (g_opt.invariants) {
      // compute_invariants (invariants_m.f90:15-198) of the same velocity
      hd += "\"Q\"\n\"C\"\n\"D\"\n\"H\"\n\"L2\"\n";
      if (!invbuf.ensure(5 * sizeof(T) * np)) {
        fail("hipMalloc failed (velocity invariants)");
        return;
      }
      launch_invariants(lg, du, (T*)invbuf.p, stream);
      auto h = host_copy((const T*)invbuf.p, 5 * np);
      if (!h) {
        fail("copy of the velocity invariants to the host failed");
        return;
      }
      for (int q = 0; q < 5; q++) j.arrays.push_back(io::Array{h->data() + q * np, np, sizeof(U) == 8});
      j.keep.push_back(h);
    }
    hd += "\"FTLE\"\nZONE\nT = \"" + ft + "\"\n";
    char es[18];
    io::format_es((double)t_np1, 18, 4, es);
    hd += "STRANDID = 0000000001, SOLUTIONTIME = " + std::string(es, 18) + "\n";
    char ijk[96];
    std::snprintf(ijk, sizeof ijk, "I = %010d J = %010d K = %010d\n", lg.n[0], lg.n[1], lg.n[2]);
    hd += ijk;
    hd += "ZONETYPE = Ordered, DATAPACKING = BLOCK\n";
    j.header = hd;
    {
      std::vector<double> f(np);
      if (hipMemcpyAsync(f.data(), dg.ftle.p, np * sizeof(double), hipMemcpyDeviceToHost, stream) != hipSuccess ||
          hipStreamSynchronize(stream) != hipSuccess) {
        fail("copy of the FTLE field to the host failed");
        return;
      }
      auto h = std::make_shared<std::vector<U>>(np);
      for (size_t a = 0; a < np; a++) (*h)[a] = (U)f[a];
      j.arrays.push_back(io::Array{h->data(), np, sizeof(U) == 8});
      j.keep.push_back(h);
    }
    wr().submit(std::move(j));
  }

  bool sync_status() {
    if (hipStreamSynchronize(stream) != hipSuccess) {
      fail("device synchronisation failed");
      return false;
    }
    harvest_timers();
    if (hipMemcpy(st_host, st, sizeof(DevStatus), hipMemcpyDeviceToHost) != hipSuccess) return false;
    oob_total = st_host->oob;
    if (st_host->err) {
      fail("device error (bcflag/tracking)");
      return false;
    }
    return true;
  }

  Diag* diag(int h) { return (h >= 1 && h <= (int)diags.size()) ? diags[h - 1].get() : nullptr; }

  // =======================================================================
  // Adjoint of the backward semi-Lagrangian FTLE over one window of
  // snapshots (the BKWD diagnostic with h = T): forward pass with flowmap
  // checkpoints every `ckpt` steps, backward pass that recomputes each
  // segment from its checkpoint (keeping every sub-cycle's flowmap) and runs
  // the sub-cycles in reverse.  The tape owns its own velocity ring, which is
  // swapped with the streaming one for the duration of each call.
  // =======================================================================
  struct Tape {
    bool have = false, fwd_done = false;
    int nsnap = 0, ckpt = 1, interp = 0, integ = 0;
    bool spl = false;
    const U *du = nullptr, *dv = nullptr, *dw = nullptr;
    std::vector<U> times;
    std::vector<int> nsub;       // per step s = 1..nsnap-1 (index s)
    U Tint = 0;
    // swapped with the streaming state while the tape runs
    DevBuf vel[2], vcoef[2];
    bool vcoef_ok[2] = {false, false};
    bool vcoef32[2] = {false, false};
    int cur = 1, vold = 0;
    U t_n = 0, t_np1 = 0;
    Blend blends[6];
    int nblend = 6;
    // state
    DevBuf fm[2], fmcoef, compfm, dxout, ftle, lam, adjfm, adjv[3], dcn, dcnp1, dun, dunp1, ovf[4];
    DevBuf tiles, oflag;         // tile-store scratch of the reverse sub-cycle scatter
    int oseq = 0;                // per-call value of its overflow flag
    int fcur = 0;
    std::vector<DevBuf> ckpts;   // fm (ghosted values) after steps 0, ckpt, 2 ckpt, ...
    std::vector<DevBuf> segf;    // per sub-cycle of a segment: the field k_sl_step read
    double t_fwd = 0, t_bwd = 0;
    size_t bytes = 0;
  };
  Tape tape;

  void swap_vel_state() {
    for (int b = 0; b < 2; b++) {
      vel[b].swap(tape.vel[b]);
      vcoef[b].swap(tape.vcoef[b]);
      std::swap(vcoef_ok[b], tape.vcoef_ok[b]);
      std::swap(vcoef32[b], tape.vcoef32[b]);
    }
    std::swap(cur, tape.cur);
    std::swap(vold, tape.vold);
    std::swap(t_n, tape.t_n);
    std::swap(t_np1, tape.t_np1);
    std::swap(nblend, tape.nblend);
    for (int b = 0; b < 6; b++) {
      blends[b].buf.swap(tape.blends[b].buf);
      std::swap(blends[b].ct, tape.blends[b].ct);
      std::swap(blends[b].storage, tape.blends[b].storage);
      std::swap(blends[b].spl, tape.blends[b].spl);
      std::swap(blends[b].valid, tape.blends[b].valid);
      std::swap(blends[b].stamp, tape.blends[b].stamp);
    }
  }

  bool ensure(DevBuf& b, size_t bytes, bool zero = false) {
    if (!b.p || b.bytes < bytes) {
      if (!get_buf(b, bytes)) {
        fail(arena ? "the allocation-free workspace is too small for this call"
                   : "hipMalloc failed (adjoint buffers; out of device memory?)");
        return false;
      }
      tape.bytes += bytes;
    }
    if (zero && !zero_async(b.p, bytes)) {
      fail("clearing a device buffer failed");
      return false;
    }
    return true;
  }

  // snapshots s-1 and s into the ring as u_n and u_np1 (ghosts, MASK zero,
  // spline coefficients); returns the CFL rate of u_np1
  // readback = false: the caller needs no CFL rate (the backward pass takes
  // the sub-cycle counts from the tape), so the step is queued without
  // waiting for the device; device errors surface at the pass's final check
  bool tape_load_step(int s, double& cfl_rate, bool readback = true) {
    const int old = cur;
    cur = 1 - cur;
    vold = old;
    vcoef_ok[cur] = false;
    vcoef_ok[old] = false;
    for (auto& b : blends) b.valid = false;
    t_n = tape.times[s - 1];
    t_np1 = tape.times[s];
    const size_t np = (size_t)g.np;
    for (int pass = 0; pass < 2; pass++) {
      const int b = pass == 0 ? old : cur;
      const int snap = pass == 0 ? s - 1 : s;
      if (!zero_async(&st->cfl_bits, sizeof(st->cfl_bits) + sizeof(st->umax_bits))) {
        fail("clearing the status words failed");
        return false;
      }
      launch_ingest<U, T>(g, f3(vel[b]), tape.du + snap * np, tape.dv + snap * np, tape.dw + snap * np, stream);
      launch_vel_ghost(g, f3(vel[b]), cf3(vel[1 - b]), st, stream);
      launch_cfl(g, f3(vel[b]), st, stream);
    }
    cfl_rate = 0;
    if (readback) {
      HIPCHK_R(hipMemcpyAsync(st_host, st, sizeof(DevStatus), hipMemcpyDeviceToHost, stream), false);
      HIPCHK_R(hipStreamSynchronize(stream), false);
      if (st_host->err) {
        fail("device error while loading snapshots");
        return false;
      }
      unsigned long long bits = st_host->cfl_bits;
      std::memcpy(&cfl_rate, &bits, sizeof(double));
    }
    if (tape.spl) {
      if (!fit_vel(cur)) return false;
      if (!fit_vel(old)) return false;
    }
    return true;
  }

  // the sub-cycle times of step s, as the forward loop accumulates them
  void step_times(int s, std::vector<U>& ts, U& dt) {
    const int ns = tape.nsub[s];
    dt = (t_np1 - t_n) / (U)ns;
    dt = (U)-1 * dt;
    ts.resize(ns);
    U t = t_np1;
    for (int k = 0; k < ns; k++) {
      ts[k] = t;
      t = t + dt;
    }
  }

  // one forward sub-cycle; keeps the field k_sl_step read in `keep` if given
  bool tape_subcycle(U t, U dt, DevBuf* keep) {
    DevBuf& a = tape.fm[tape.fcur];
    DevBuf& b = tape.fm[1 - tape.fcur];
    CF3<T> src = cf3(a);
    const size_t fb = 3 * fbytes();
    if (tape.spl) {
      // the coefficients the step reads, fitted straight into the kept
      // field when the sub-cycle's field is kept
      DevBuf& cb = keep ? *keep : tape.fmcoef;
      if (!ensure(cb, fb)) return false;
      launch_spline_fit(g, f3(cb), 3, stream, src.c);
      src = cf3(cb);
    } else if (keep) {
      if (!ensure(*keep, fb)) return false;
      HIPCHK_R(hipMemcpyAsync(keep->p, a.p, fb, hipMemcpyDeviceToDevice, stream), false);
    }
    const StageParams<T> sp = stage_params(t, dt, tape.spl);
    launch_sl_step(g, tape.interp, sp, src, f3(b), st, stream, g_opt.sl_tiled);
    launch_fm_ghost(g, f3(b), cf3(a), st, stream);
    tape.fcur = 1 - tape.fcur;
    return true;
  }

  // arguments of a forward window call (both entry points)
  bool window_args_ok(int nsnap, const U* du, const U* dv, const U* dw, const U* times, int ckpt, const char* who) {
    const int interp = g_opt.interpolator;
    if (!hg.rect) {
      fail("the adjoint pair on a non-rectilinear grid is not supported by this build");
      return false;
    }
    if (nsnap < 2 || !du || !dv || !dw || !times || ckpt < 1) {
      fail(std::string(who) + ": need at least 2 snapshots, device arrays, times and checkpoint >= 1");
      return false;
    }
    if (g_opt.vel_storage != kStoreNative) {
      fail("the adjoint runs with VELOCITY_STORAGE equal to the compute precision");
      return false;
    }
    if (interp != kNearest && interp != kLinear && interp != kQuadratic && interp != kCubic) {
      fail("the adjoint supports INTERPOLATOR NEAREST_NBR, LINEAR, QUADRATIC and CUBIC");
      return false;
    }
    for (int s = 1; s < nsnap; s++)
      if (!(times[s] > times[s - 1])) {
        fail(std::string(who) + ": snapshot times must increase");
        return false;
      }
    return true;
  }

  int adjoint_forward(int nsnap, const U* du, const U* dv, const U* dw, const U* times, int ckpt,
                      double* d_ftle) override {
    if (!window_args_ok(nsnap, du, dv, dw, times, ckpt, "cfd2lcs_adjoint_forward_c")) return 1;
    return window_forward(nsnap, du, dv, dw, times, ckpt, d_ftle, true);
  }

  // the BKWD FTLE of one window of device snapshots (the diagnostic with
  // h = T); keep: record the tape (checkpoints every ckpt steps) for
  // adjoint_backward, else write the FTLE straight to d_ftle and keep nothing
  int window_forward(int nsnap, const U* du, const U* dv, const U* dw, const U* times, int ckpt, double* d_ftle,
                     bool keep) {
    const int interp = g_opt.interpolator;
    auto t0 = std::chrono::steady_clock::now();
    tape.have = false;
    tape.fwd_done = false;
    tape.nsnap = nsnap;
    tape.ckpt = ckpt;
    tape.interp = interp;
    tape.integ = g_opt.integrator;
    tape.spl = is_spline(interp);
    tape.du = du; tape.dv = dv; tape.dw = dw;
    tape.times.assign(times, times + nsnap);
    tape.nsub.assign(nsnap, 0);
    tape.Tint = times[nsnap - 1] - times[0];
    if (tape.spl && !ensure_spline(interp)) return 1;
    const size_t fb = 3 * fbytes();
    for (int b = 0; b < 2; b++)
      if (!ensure(tape.vel[b], fb, true) || !ensure(tape.fm[b], fb, true)) return 1;
    tape.fcur = 0;
    const int nseg = keep ? (nsnap - 2) / ckpt + 1 : 0;
    tape.ckpts.resize(nseg);
    swap_vel_state();
    bool ok = true;
    for (int s = 1; s < nsnap && ok; s++) {
      if (keep && (s - 1) % ckpt == 0) {
        DevBuf& c = tape.ckpts[(s - 1) / ckpt];
        ok = ensure(c, fb);
        if (ok) ok = hipMemcpyAsync(c.p, tape.fm[tape.fcur].p, fb, hipMemcpyDeviceToDevice, stream) == hipSuccess;
        if (!ok) break;
      }
      double cfl_rate = 0;
      if (!(ok = tape_load_step(s, cfl_rate))) break;
      const U dt_f = t_np1 - t_n;
      const U cmax = (U)((T)cfl_rate) * dt_f;
      const double r = std::ceil((double)(cmax / g_opt.cfl_max));
      if (r > 1e6) {
        fail("CFL sub-cycle count above 1e6");
        ok = false;
        break;
      }
      tape.nsub[s] = r >= 1 ? (int)r : 1;
      std::vector<U> ts;
      U dt;
      step_times(s, ts, dt);
      for (int k = 0; k < tape.nsub[s] && ok; k++) ok = tape_subcycle(ts[k], dt, nullptr);
    }
    swap_vel_state();
    if (!ok) return 1;
    // FTLE of the window (the composition of a single slot is the slot itself)
    if (!ensure(tape.dxout, 3 * sizeof(T) * (size_t)g.np) || !ensure(tape.compfm, fb, true) ||
        (keep && !ensure(tape.ftle, sizeof(double) * (size_t)g.np)))
      return 1;
    double* fout = keep ? (double*)tape.ftle.p : d_ftle;
    launch_interior(g, cf3(tape.fm[tape.fcur]), p3(tape.dxout), stream);
    launch_field_from_origin(g, cp3(tape.dxout), f3(tape.compfm), stream);
    const double tiny = sizeof(U) == 8 ? DBL_MIN : (double)FLT_MIN;
    launch_ftle(g, cf3(tape.compfm), (double)tape.Tint, g_opt.incompressible, tiny, fout, p3(tape.dxout), stream);
    if (keep && d_ftle)
      HIPCHK_R(hipMemcpyAsync(d_ftle, tape.ftle.p, sizeof(double) * (size_t)g.np, hipMemcpyDeviceToDevice, stream),
               1);
    HIPCHK_R(hipGetLastError(), 1);
    HIPCHK_R(hipMemcpyAsync(st_host, st, sizeof(DevStatus), hipMemcpyDeviceToHost, stream), 1);
    HIPCHK_R(hipStreamSynchronize(stream), 1);
    if (st_host->err) {
      fail("device error in the adjoint forward pass");
      return 1;
    }
    tape.have = keep;
    tape.fwd_done = keep;
    tape.t_fwd = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
  }


  // cfd2lcs_adjoint_forward_noalloc_c: the FTLE of cfd2lcs_adjoint_forward_c
  // with every buffer carved from the fixed workspace of scratch.hip instead
  // of allocated, and no tape kept.  The engine's own buffers (velocity ring,
  // blend pool, status words, a tape of an earlier adjoint_forward) are set
  // aside for the call and restored after it; spline factors already set up
  // are reused, missing ones live in the workspace for the call only.
  int adjoint_forward_noalloc(int nsnap, const U* du, const U* dv, const U* dw, const U* times,
                              double* d_ftle) override {
    if (!window_args_ok(nsnap, du, dv, dw, times, 1, "cfd2lcs_adjoint_forward_noalloc_c")) return 1;
    if (!d_ftle) {
      fail("cfd2lcs_adjoint_forward_noalloc_c: null FTLE pointer");
      return 1;
    }
    void* ws = ws_ok ? scratch_ptr() : nullptr;
    if (!ws) {
      fail("cfd2lcs_adjoint_forward_noalloc_c: the device workspace was not set up on this device");
      return 1;
    }
    const int interp = g_opt.interpolator;
    const bool spl = is_spline(interp);
    // workspace plan: status words, u_n / u_np1 (+ coefficients), three
    // stage blends, the flowmap pair (+ coefficients), the composed field,
    // the interior displacement, and spline factors if not yet set up
    const size_t fb = 3 * fbytes();
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    size_t need = al(sizeof(DevStatus)) + (2 + (spl ? 2 : 0) + 3 + 2 + (spl ? 1 : 0) + 1) * al(fb) +
                  al(3 * sizeof(T) * (size_t)g.np);
    const bool own_factors = spl && spline_k != spline_order_for(interp);
    if (own_factors)
      for (int d = 0; d < 3; d++) {
        // the same sizes grid_factors uploads
        SplineAxis ax;
        const int kk = std::max(std::min(spline_order_for(interp), hg.N[d] - 1), 2);
        if (!spline_axis(hg.xl[d], kk, ax).empty()) {
          fail("cfd2lcs_adjoint_forward_noalloc_c: spline set-up failed");
          return 1;
        }
        need += al(sizeof(T) * (ax.knots.size() + ax.band.size() + ax.poly.size()));
      }
    if (need > kScratchBytes) {
      fail("cfd2lcs_adjoint_forward_noalloc_c: this grid needs " + std::to_string(need >> 20) +
           " MiB of workspace, more than the fixed " + std::to_string(kScratchBytes >> 20) + " MiB");
      return 1;
    }
    Arena ar;
    ar.base = (char*)ws;
    ar.size = kScratchBytes;
    // set aside what the shared code would touch
    Tape saved = std::move(tape);
    tape = Tape();
    DevStatus* st_saved = st;
    DevBuf sm_saved[3];
    GridDev<T> g_saved = g;
    const int k_saved = spline_k;
    const bool vc_saved[2] = {vcoef_ok[0], vcoef_ok[1]};
    arena = &ar;
    bool ok = true;
    DevBuf stbuf;
    ok = get_buf(stbuf, sizeof(DevStatus));
    if (ok) {
      st = (DevStatus*)stbuf.p;
      ok = zero_async(st, sizeof(DevStatus));
    }
    for (int b = 0; b < 3 && ok; b++) ok = get_buf(tape.blends[b].buf, fb);
    for (int b = 0; b < 2 && ok; b++) ok = get_buf(tape.vel[b], fb) && (!spl || get_buf(tape.vcoef[b], fb));
    if (ok && own_factors) {
      for (int d = 0; d < 3; d++) sm_saved[d].swap(spline_mem[d]);
      spline_k = 0;
    }
    int rc = 1;
    if (ok) rc = window_forward(nsnap, du, dv, dw, times, 1, d_ftle, false);
    else fail("cfd2lcs_adjoint_forward_noalloc_c: workspace plan exceeded");
    // everything above ran on the stream; drain it before the workspace and
    // the status words are handed back
    (void)hipStreamSynchronize(stream);
    arena = nullptr;
    st = st_saved;
    if (own_factors) {
      for (int d = 0; d < 3; d++) sm_saved[d].swap(spline_mem[d]);
      g.kord[0] = g_saved.kord[0]; g.kord[1] = g_saved.kord[1]; g.kord[2] = g_saved.kord[2];
      for (int d = 0; d < 3; d++) {
        g.kn[d] = g_saved.kn[d];
        g.band[d] = g_saved.band[d];
        g.bpoly[d] = g_saved.bpoly[d];
      }
      spline_k = k_saved;
      vcoef_ok[0] = vc_saved[0];
      vcoef_ok[1] = vc_saved[1];
    }
    tape = std::move(saved);
    return rc;
  }

  int adjoint_backward(const double* d_dftle, U* gu, U* gv, U* gw) override {
    if (!tape.have) {
      fail("cfd2lcs_adjoint_backward_c called without a successful cfd2lcs_adjoint_forward_c");
      return 1;
    }
    if (!d_dftle || !gu || !gv || !gw) {
      fail("cfd2lcs_adjoint_backward_c: null device pointer");
      return 1;
    }
    auto t0 = std::chrono::steady_clock::now();
    const size_t fb = 3 * fbytes(), np = (size_t)g.np;
    const int nsnap = tape.nsnap;
    for (DevBuf* b : {&tape.lam, &tape.adjfm, &tape.adjv[0], &tape.adjv[1], &tape.adjv[2], &tape.dcn, &tape.dcnp1,
                      &tape.dun, &tape.dunp1, &tape.ovf[0], &tape.ovf[1], &tape.ovf[2], &tape.ovf[3]})
      if (!ensure(*b, fb, true)) return 1;
    for (U* o : {gu, gv, gw}) HIPCHK_R(hipMemsetAsync(o, 0, sizeof(U) * np * nsnap, stream), 1);
    const double tiny = sizeof(U) == 8 ? DBL_MIN : (double)FLT_MIN;
    // FTLE -> flowmap of the window's end
    launch_ftle_adjoint(g, cf3(tape.compfm), (double)tape.Tint, g_opt.incompressible, tiny, d_dftle, f3(tape.lam),
                        stream);
    launch_fold_ghost(g, f3(tape.lam), 1, stream);
    int maxsub = 0;
    for (int s = 1; s < nsnap; s++) maxsub = std::max(maxsub, tape.nsub[s]);
    swap_vel_state();
    bool ok = true;
    const int nseg = (int)tape.ckpts.size();
    for (int seg = nseg - 1; seg >= 0 && ok; seg--) {
      const int s0 = seg * tape.ckpt + 1, s1 = std::min(s0 + tape.ckpt - 1, nsnap - 1);
      // recompute the segment from its checkpoint, keeping each sub-cycle's
      // flowmap as k_sl_step read it
      ok = hipMemcpyAsync(tape.fm[tape.fcur].p, tape.ckpts[seg].p, fb, hipMemcpyDeviceToDevice, stream) ==
           hipSuccess;
      const size_t need = (size_t)(s1 - s0 + 1) * maxsub;
      if (tape.segf.size() < need) tape.segf.resize(need);
      for (int s = s0; s <= s1 && ok; s++) {
        double cfl_rate;
        if (!(ok = tape_load_step(s, cfl_rate, false))) break;
        std::vector<U> ts;
        U dt;
        step_times(s, ts, dt);
        for (int k = 0; k < tape.nsub[s] && ok; k++)
          ok = tape_subcycle(ts[k], dt, &tape.segf[(size_t)(s - s0) * maxsub + k]);
      }
      // reverse sweep over the segment
      for (int s = s1; s >= s0 && ok; s--) {
        double cfl_rate;
        if (!(ok = tape_load_step(s, cfl_rate, false))) break;
        std::vector<U> ts;
        U dt;
        step_times(s, ts, dt);
        for (DevBuf* b : {&tape.dcn, &tape.dcnp1, &tape.dun, &tape.dunp1})
          HIPCHK_R(hipMemsetAsync(b->p, 0, fb, stream), 1);
        for (int k = tape.nsub[s] - 1; k >= 0 && ok; k--) {
          const StageParams<T> sp = stage_params(ts[k], dt, tape.spl);
          // stage ct values and their adjoint buffers (stages >= 2)
          const U half = (U)1 / (U)2;
          U c[4] = {0, 0, 0, 0};
          int nst = 1;
          switch (tape.integ) {
            case kRK2: c[1] = ctf(ts[k] + half * dt, t_n, t_np1); nst = 2; break;
            case kRK3:
              c[1] = ctf(ts[k] + half * dt, t_n, t_np1);
              c[2] = ctf(ts[k] + dt, t_n, t_np1);
              nst = 3;
              break;
            case kRK4:
              c[1] = ctf(ts[k] + half * dt, t_n, t_np1);
              c[2] = c[1];
              c[3] = ctf(ts[k] + dt, t_n, t_np1);
              nst = 4;
              break;
            default: break;
          }
          int sbuf[4] = {0, 0, 0, 0};
          U bct[3] = {0, 0, 0};
          int nb = 0;
          for (int a = 1; a < nst; a++) {
            int f = -1;
            for (int b = 0; b < nb; b++)
              if (bct[b] == c[a]) f = b;
            if (f < 0) {
              f = nb++;
              bct[f] = c[a];
            }
            sbuf[a] = f;
          }
          const DevBuf& fread = tape.segf[(size_t)(s - s0) * maxsub + k];
          F3<T> av[3] = {f3(tape.adjv[0]), f3(tape.adjv[1]), f3(tape.adjv[2])};
          F3<T> ov[4] = {f3(tape.ovf[0]), f3(tape.ovf[1]), f3(tape.ovf[2]), f3(tape.ovf[3])};
          // the scatter stores each brick's tile and gathers them in one
          // pass (8 tiles per node and phase, at most 3 phases) that also
          // unblends the stage buffers; above kTileStoreMax bytes it adds the
          // tiles in place, colour by colour, and the buffers are unblended
          // below
          TileStore<T> tstore{nullptr, nullptr, 0};
          const size_t tb = sl_adjoint_tile_bytes(g, tape.interp, nb);
          if (tb <= kTileStoreMax && !g_opt.adj_colours) {
            if (!tape.oflag.p) {
              // -1: a value no call number takes
              if (!ensure(tape.oflag, sizeof(int))) return 1;
              HIPCHK_R(hipMemsetAsync(tape.oflag.p, 0xFF, sizeof(int), stream), 1);
            }
            if (!ensure(tape.tiles, tb)) return 1;
            tstore = TileStore<T>{(T*)tape.tiles.p, (int*)tape.oflag.p, tape.oseq};
            tape.oseq = (tape.oseq + 1) & 0x7FFFFFFF;
          } else {
            HIPCHK_R(hipMemsetAsync(tape.adjfm.p, 0, fb, stream), 1);
          }
          const T bc[3] = {(T)bct[0], (T)bct[1], (T)bct[2]};
          const bool unblended =
              launch_sl_adjoint(g, tape.interp, sp, cf3(const_cast<DevBuf&>(fread)), cf3(tape.lam), f3(tape.adjfm),
                                av, ov, sbuf, nb, f3(tape.dun), f3(tape.dunp1), tstore.tiles ? &tstore : nullptr,
                                bc, f3(tape.dcn), f3(tape.dcnp1), stream);
          for (int b = 0; b < nb && !unblended; b++)
            launch_adj_unblend(g, f3(tape.adjv[b]), (T)bct[b], f3(tape.dcn), f3(tape.dcnp1), stream);
          if (tape.spl) launch_spline_fit_T(g, f3(tape.adjfm), stream);
          launch_fold_ghost(g, f3(tape.adjfm), 1, stream);
          tape.lam.swap(tape.adjfm);
        }
        // velocity adjoints of this step, back to the snapshots
        if (tape.spl) {
          launch_spline_fit_T(g, f3(tape.dcn), stream);
          launch_spline_fit_T(g, f3(tape.dcnp1), stream);
        }
        launch_adj_unblend(g, f3(tape.dun), (T)0, f3(tape.dcn), f3(tape.dcnp1), stream);
        launch_adj_unblend(g, f3(tape.dunp1), (T)1, f3(tape.dcn), f3(tape.dcnp1), stream);
        launch_fold_ghost(g, f3(tape.dcn), 0, stream);
        launch_fold_ghost(g, f3(tape.dcnp1), 0, stream);
        launch_adj_to_user<T, U>(g, cf3(tape.dcn), gu + (s - 1) * np, gv + (s - 1) * np, gw + (s - 1) * np, 1,
                                 stream);
        launch_adj_to_user<T, U>(g, cf3(tape.dcnp1), gu + s * np, gv + s * np, gw + s * np, 1, stream);
        ok = hipGetLastError() == hipSuccess;
      }
    }
    swap_vel_state();
    if (!ok) {
      if (!g_error) fail("adjoint backward pass failed");
      return 1;
    }
    HIPCHK_R(hipMemcpyAsync(st_host, st, sizeof(DevStatus), hipMemcpyDeviceToHost, stream), 1);
    HIPCHK_R(hipStreamSynchronize(stream), 1);
    if (st_host->err) {
      fail("device error in the adjoint backward pass");
      return 1;
    }
    tape.t_bwd = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
  }

  void adjoint_release() override {
    (void)hipStreamSynchronize(stream);
    tape = Tape();
  }

  void adjoint_stats(double out[4]) override {
    out[0] = tape.t_fwd;
    out[1] = tape.t_bwd;
    out[2] = (double)tape.bytes;
    long long tot = 0;
    for (int v : tape.nsub) tot += v;
    out[3] = (double)tot;
  }

  int copy_out(CF3<T> src, int ghosted, U* a, U* b, U* c) { return copy_out(g, src, ghosted, a, b, c); }
  int copy_out(const GridDev<T>& gd, CF3<T> src, int ghosted, U* a, U* b, U* c) {
    const long long np = gd.np;
    if (!stage_out.ensure(3 * sizeof(U) * (size_t)np)) return 1;
    U* o = (U*)stage_out.p;
    launch_interior_to_user<T, U>(gd, src, ghosted, o, o + np, o + 2 * np, stream);
    const size_t b1 = sizeof(U) * (size_t)np;
    if (hipMemcpyAsync(a, o, b1, hipMemcpyDeviceToHost, stream) != hipSuccess) return 1;
    if (b && hipMemcpyAsync(b, o + np, b1, hipMemcpyDeviceToHost, stream) != hipSuccess) return 1;
    if (c && hipMemcpyAsync(c, o + 2 * np, b1, hipMemcpyDeviceToHost, stream) != hipSuccess) return 1;
    return hipStreamSynchronize(stream) == hipSuccess ? 0 : 1;
  }

  int get_flowmap(int h, U* a, U* b, U* c) override {
    Diag* dg = diag(h);
    if (!dg) return 1;
    if (dg->dir > 0) return copy_out(G(*dg), fwd_dx_by_origin(*dg), 0, a, b, c);
    if (dg->fm32) {
      // widened into a temporary ghosted field first
      DevBuf w;
      if (!w.alloc(3 * fbytes(NT(*dg)))) return 1;
      launch_widen((const float*)dg->fm[dg->fcur].p, (T*)w.p, 3 * NT(*dg), stream);
      return copy_out(G(*dg), cf3(w, NT(*dg)), 1, a, b, c);
    }
    return copy_out(G(*dg), cf3(dg->fm[dg->fcur], NT(*dg)), 1, a, b, c);
  }
  int get_composed(int h, U* a, U* b, U* c) override {
    Diag* dg = diag(h);
    if (!dg) return 1;
    return copy_out(G(*dg), cp3(dg->comp, NP(*dg)), 0, a, b, c);
  }
  int get_ftle(int h, U* out) override {
    Diag* dg = diag(h);
    if (!dg) return 1;
    const long long np = NP(*dg);
    std::vector<double> tmp(np);
    if (hipStreamSynchronize(stream) != hipSuccess) return 1;
    if (hipMemcpy(tmp.data(), dg->ftle.p, sizeof(double) * np, hipMemcpyDeviceToHost) != hipSuccess) return 1;
    for (long long p = 0; p < np; p++) out[p] = (U)tmp[p];
    return 0;
  }
  // extents and (when x is given) node coordinates of a diagnostic's grid
  int get_grid(int h, int n[3], U* x, U* y, U* z) override {
    Diag* dg = diag(h);
    if (!dg) return 1;
    const HostGrid& hgd = HG(*dg);
    for (int d = 0; d < 3; d++) n[d] = hgd.n[d];
    U* o[3] = {x, y, z};
    for (int k = 1; k <= hgd.n[2]; k++)
      for (int j = 1; j <= hgd.n[1]; j++)
        for (int i = 1; i <= hgd.n[0]; i++)
          for (int d = 0; d < 3; d++)
            if (o[d]) o[d][hgd.pid(i, j, k)] = (U)hgd.g[d][hgd.id(i, j, k)];
    return 0;
  }
  const double* ftle_device(int h) override {
    Diag* dg = diag(h);
    if (!dg) return nullptr;
    (void)hipStreamSynchronize(stream);
    return (const double*)dg->ftle.p;
  }
  int noutput(int h) override {
    Diag* dg = diag(h);
    return dg ? dg->noutput : -1;
  }
  void stats(double out[8]) override {
    sync_status();
    out[0] = nsub_last;
    out[1] = (double)integ_fwd;
    out[2] = (double)integ_bkwd;
    out[3] = tsum[0];
    out[4] = tsum[1];
    out[5] = tsum[2];
    out[6] = tsum[3];
    out[7] = (double)oob_total;
  }
  int debug_grid(double* lines, int* flag, double* bn, double* ps, double* bb) override {
    if (lines) {
      size_t o = 0;
      for (int d = 0; d < 3; d++)
        for (int a = 0; a < hg.N[d]; a++) lines[o++] = hg.line[d][a];
    }
    if (flag)
      for (long long l = 0; l < hg.nt; l++) flag[l] = hg.flag[l];
    if (bn)
      for (int d = 0; d < 3; d++)
        for (long long p = 0; p < hg.np; p++) bn[d * hg.np + p] = hg.bn[d][p];
    if (ps)
      for (int c = 0; c < 27; c++)
        for (int d = 0; d < 3; d++) ps[3 * c + d] = hg.ps[c][d];
    if (bb)
      for (int d = 0; d < 6; d++) bb[d] = hg.bb[d];
    return 0;
  }
};

// ---------------------------------------------------------------------------
// library-global state
// ---------------------------------------------------------------------------
static std::unique_ptr<HostGrid> g_grid;
static std::unique_ptr<EngineBase> g_eng;
static int g_engine_prec = 0;

static bool ensure_engine() {
  if (g_eng) return true;
  if (!g_grid) {
    fail("cfd2lcs_init_c has not been called");
    return false;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    fail("no HIP device visible: the FTLE path runs on the GPU only");
    return false;
  }
  if (g_opt.precision == LCS_FP64) {
    auto e = std::make_unique<Engine<double>>();
    if (!e->init(*g_grid)) return false;
    g_eng = std::move(e);
  } else {
    auto e = std::make_unique<Engine<float>>();
    if (!e->init(*g_grid)) return false;
    g_eng = std::move(e);
  }
  g_engine_prec = g_opt.precision;
  return true;
}

}  // namespace lcs

using namespace lcs;

static std::string trim_label(const char* s, size_t n) {
  std::string r(s, n);
  while (!r.empty() && (r.back() == ' ' || r.back() == '\0')) r.pop_back();
  return r;
}

extern "C" {

void cfd2lcs_init_c(MPI_Comm usercomm, int n[3], int offset[3], lcsdata_t* x, lcsdata_t* y, lcsdata_t* z,
                    int* flag) {
  (void)usercomm;
  (void)offset;  // one rank: the minimum offset is normalised to 0 (sgrid_m.f90:90-93)
  if (g_error) return;
  g_eng.reset();
  const long long np = (long long)n[0] * n[1] * n[2];
  if (n[0] < 1 || n[1] < 1 || n[2] < 1) {
    fail("ng > ni, nj or nk");
    return;
  }
  std::vector<double> X(np), Y(np), Z(np);
  for (long long p = 0; p < np; p++) {
    X[p] = x[p];
    Y[p] = y[p];
    Z[p] = z[p];
  }
  auto hg = std::make_unique<HostGrid>();
  // ghost edges may depart from the coordinate lines by the rounding of the
  // user's coordinates (lcsdata_t)
  std::string e = hg->build(n, X.data(), Y.data(), Z.data(), flag, sizeof(lcsdata_t) == 8 ? 1e-9 : 1e-6);
  if (!e.empty()) {
    fail(e);
    return;
  }
  g_grid = std::move(hg);
  ensure_engine();
}

int cfd2lcs_diagnostic_init_c(int lcs_type, int resolution, lcsdata_t t, lcsdata_t h, char label[]) {
  if (g_error) return 0;
  if (!ensure_engine()) return 0;
  if (lcs_type == LP_TRACER) {
    fail("LP_TRACER diagnostics are not supported by this build");
    return 0;
  }
  if (lcs_type != FTLE_FWD && lcs_type != FTLE_BKWD) {
    fail("bad specification for lcs_type. Options are: FTLE_FWD, FTLE_BKWD, LP_TRACER");
    return 0;
  }
  if (g_eng->ndiag() >= 1000) {
    fail("NLCS exceeds NMAX_STRUCT");
    return 0;
  }
  return g_eng->add_diag(lcs_type, resolution, t, h, label);
}

void cfd2lcs_update_c(int n[3], lcsdata_t* u, lcsdata_t* v, lcsdata_t* w, lcsdata_t time) {
  if (g_error) return;
  if (!ensure_engine()) return;
  if (n[0] != g_grid->n[0] || n[1] != g_grid->n[1] || n[2] != g_grid->n[2]) {
    fail("received velocity array of incorrect dimension");
    return;
  }
  g_eng->update(u, v, w, false, time);
}

void cfd2lcs_update_device_c(int n[3], const lcsdata_t* d_u, const lcsdata_t* d_v, const lcsdata_t* d_w,
                             lcsdata_t time) {
  if (g_error) return;
  if (!ensure_engine()) return;
  if (n[0] != g_grid->n[0] || n[1] != g_grid->n[1] || n[2] != g_grid->n[2]) {
    fail("received velocity array of incorrect dimension");
    return;
  }
  g_eng->update(d_u, d_v, d_w, true, time);
}

void cfd2lcs_diagnostic_destroy_c(int lcs_handle) {
  // the reference leaves this as a TODO (cfd2lcs.f90:372-379): nothing is freed
  (void)lcs_handle;
}

void cfd2lcs_set_option_c(char option[], int val) {
  if (g_error) return;
  const std::string o = option ? option : "";
  bool warn = false;
  auto boolopt = [&](int& dst) {
    if (val == LCS_TRUE) dst = 1;
    else if (val == LCS_FALSE) dst = 0;
    else warn = true;
  };
  if (o == "SYNCTIMER") boolopt(g_opt.sync_timer);
  else if (o == "DEBUG") boolopt(g_opt.verbose);
  else if (o == "INCOMPRESSIBLE") boolopt(g_opt.incompressible);
  else if (o == "WRITE_FLOWMAP") boolopt(g_opt.flowmap_io);
  else if (o == "WRITE_VELOCITY") boolopt(g_opt.velocity_io);
  else if (o == "WRITE_BCFLAG") boolopt(g_opt.bcflag_io);
  else if (o == "WRITE_LCS") boolopt(g_opt.write_lcs);
  else if (o == "WRITE_SLOTS") boolopt(g_opt.write_slots);
  else if (o == "VELOCITY_INVARIANTS") boolopt(g_opt.invariants);
  else if (o == "AUX_GRID") boolopt(g_opt.aux_grid);
  else if (o == "ADJOINT_COLOURS") boolopt(g_opt.adj_colours);
  else if (o == "KERNEL_TIMING") {
    boolopt(g_opt.kernel_timing);
  } else if (o == "VELOCITY_STORAGE") {
    // storage of the gathered stage velocities: the compute precision,
    // 16-bit (FP32 compute only) or FP32 under FP64 compute
    if (val == LCS_FP16) g_opt.vel_storage = kStoreF16;
    else if (val == LCS_BF16) g_opt.vel_storage = kStoreBF16;
    else if (val == LCS_FP32) g_opt.vel_storage = sizeof(U) == 8 ? kStoreF32 : kStoreNative;
    else if (val == LCS_FP64) g_opt.vel_storage = kStoreNative;
    else warn = true;
  } else if (o == "FLOWMAP_STORAGE") {
    // storage of the backward flowmap from the next h-interval reset on: the
    // compute precision, or FP32 under FP64 compute
    if (val == LCS_FP32) g_opt.fm_storage = sizeof(U) == 8 ? kStoreF32 : kStoreNative;
    else if (val == LCS_FP64) g_opt.fm_storage = kStoreNative;
    else warn = true;
  } else if (o == "COMPENSATED") {
    boolopt(g_opt.compensated);
  } else if (o == "SL_TILED") {
    boolopt(g_opt.sl_tiled);
  } else if (o == "PARTICLE_SORT") {
    if (val < 0) warn = true;
    else g_opt.particle_sort = val;
  } else if (o == "COMPOSE_SORT") {
    if (val < 0) warn = true;
    else g_opt.compose_sort = val;
  } else if (o == "SPLINE_LDS") {
    if (val != LCS_TRUE && val != LCS_FALSE) warn = true;
    else g_opt.spline_lds = val;
  } else if (o == "UPDATE_FREQ") {
  } else if (o == "INTERPOLATOR") {
    if (val != NEAREST_NBR && val != LINEAR && val != QUADRATIC && val != CUBIC && val != TSE && val != TSE_LIMIT)
      warn = true;
    g_opt.interpolator = val;
  } else if (o == "INTEGRATOR") {
    if (val < EULER || val > RK4) warn = true;
    g_opt.integrator = val;
  } else if (o == "PRECISION") {
    // device compute precision; switching rebuilds the device state, which
    // is allowed before the first diagnostic_init / update only
    if (val != LCS_FP64 && val != LCS_FP32) {
      warn = true;
    } else {
      g_opt.precision = val;
      if (g_eng && g_engine_prec != val) {
        if (!g_eng->fresh()) {
          fail("PRECISION must be set before cfd2lcs_diagnostic_init_c and cfd2lcs_update_c");
          return;
        }
        g_eng.reset();
        ensure_engine();
      }
    }
  } else {
    warn = true;
  }
  if (warn) g_opt.warn_opt = true;
  if (g_opt.warn_opt) fail("bad libcfd2lcs option/val combination: " + o + "," + std::to_string(val));
}

void cfd2lcs_set_param_c(char param[], lcsdata_t val) {
  if (g_error) return;
  const std::string p = param ? param : "";
  bool warn = false;
  if (p == "CFL") g_opt.cfl_max = val;
  else if (p == "TRACER_INJECT_X" || p == "TRACER_INJECT_Y" || p == "TRACER_INJECT_Z" || p == "TRACER_INJECT_RADIUS") {
  } else warn = true;
  if (warn) g_opt.warn_par = true;
  if (g_opt.warn_par) fail("unknown libcfd2lcs parameter: " + p);
}

void cfd2lcs_finalize_c(void) {
  g_eng.reset();
  g_grid.reset();
  g_opt = Options();
  g_error = 0;
  g_engine_prec = 0;
}

int cfd2lcs_error_c(void) { return g_error; }

int cfd2lcs_flush_output_c(void) {
  if (!g_eng) return g_error ? -1 : 0;
  return g_eng->flush_output();
}

int cfd2lcs_format_es_c(const double* v, long long n, int w, int d, char* out) {
  if (!v || !out || n < 0 || w < 1 || w > 64 || d < 0 || d > 30) return -1;
  for (long long a = 0; a < n; a++) {
    if (w == 14 && d == 6) io::format_es14_6(v[a], out + a * w);
    else io::format_es(v[a], w, d, out + a * w);
  }
  return 0;
}

int cfd2lcs_get_flowmap_c(int lcs_handle, lcsdata_t* fx, lcsdata_t* fy, lcsdata_t* fz) {
  if (!g_eng) return 1;
  return g_eng->get_flowmap(lcs_handle, fx, fy, fz);
}
int cfd2lcs_get_composed_c(int lcs_handle, lcsdata_t* fx, lcsdata_t* fy, lcsdata_t* fz) {
  if (!g_eng) return 1;
  return g_eng->get_composed(lcs_handle, fx, fy, fz);
}
int cfd2lcs_get_ftle_c(int lcs_handle, lcsdata_t* ftle) {
  if (!g_eng) return 1;
  return g_eng->get_ftle(lcs_handle, ftle);
}
int cfd2lcs_get_lcs_grid_c(int lcs_handle, int n[3], lcsdata_t* x, lcsdata_t* y, lcsdata_t* z) {
  if (!g_eng) return 1;
  return g_eng->get_grid(lcs_handle, n, x, y, z);
}
const double* cfd2lcs_get_ftle_device_c(int lcs_handle) { return g_eng ? g_eng->ftle_device(lcs_handle) : nullptr; }
int cfd2lcs_get_noutput_c(int lcs_handle) { return g_eng ? g_eng->noutput(lcs_handle) : -1; }
void cfd2lcs_get_stats_c(double out[8]) {
  if (!g_eng) {
    for (int a = 0; a < 8; a++) out[a] = 0;
    return;
  }
  g_eng->stats(out);
}

int cfd2lcs_get_timers_c(double* seconds, long long* counts, int n) {
  if (!g_eng) return 0;
  return g_eng->timers(seconds, counts, n);
}

// ---- adjoint of the backward semi-Lagrangian FTLE (no reference counterpart)
int cfd2lcs_adjoint_forward_c(int n[3], int nsnap, const lcsdata_t* d_u, const lcsdata_t* d_v,
                              const lcsdata_t* d_w, const lcsdata_t* times, int checkpoint_every, double* d_ftle) {
  if (g_error) return 1;
  if (!ensure_engine()) return 1;
  if (!n || n[0] != g_grid->n[0] || n[1] != g_grid->n[1] || n[2] != g_grid->n[2]) {
    fail("received velocity array of incorrect dimension");
    return 1;
  }
  return g_eng->adjoint_forward(nsnap, d_u, d_v, d_w, times, checkpoint_every, d_ftle);
}

int cfd2lcs_adjoint_forward_noalloc_c(int n[3], int nsnap, const lcsdata_t* d_u, const lcsdata_t* d_v,
                                      const lcsdata_t* d_w, const lcsdata_t* times, double* d_ftle) {
  if (g_error) return 1;
  // the engine (device grid, stream, status words) exists from
  // cfd2lcs_init_c on; this entry point never creates it
  if (!g_eng) {
    fail("cfd2lcs_adjoint_forward_noalloc_c: call cfd2lcs_init_c first");
    return 1;
  }
  if (!n || n[0] != g_grid->n[0] || n[1] != g_grid->n[1] || n[2] != g_grid->n[2]) {
    fail("received velocity array of incorrect dimension");
    return 1;
  }
  return g_eng->adjoint_forward_noalloc(nsnap, d_u, d_v, d_w, times, d_ftle);
}

int cfd2lcs_adjoint_backward_c(const double* d_dftle, lcsdata_t* d_gu, lcsdata_t* d_gv, lcsdata_t* d_gw) {
  if (g_error) return 1;
  if (!g_eng) {
    fail("cfd2lcs_adjoint_backward_c called before cfd2lcs_init_c");
    return 1;
  }
  return g_eng->adjoint_backward(d_dftle, d_gu, d_gv, d_gw);
}

void cfd2lcs_adjoint_release_c(void) {
  if (g_eng) g_eng->adjoint_release();
}

void cfd2lcs_adjoint_stats_c(double out[4]) {
  for (int a = 0; a < 4; a++) out[a] = 0;
  if (g_eng) g_eng->adjoint_stats(out);
}

// host-side grid description (for tests): lines[N0+N1+N2], flag[nt],
// bn[3*np], ps[81], bb[6]; any pointer may be NULL
int cfd2lcs_debug_grid_c(double* lines, int* flag, double* bn, double* ps, double* bb) {
  if (!g_eng) return 1;
  return g_eng->debug_grid(lines, flag, bn, ps, bb);
}

// ---- Fortran-callable entry points (the reference's module procedures as
// seen from C, cfd2lcs_c.c:16-58); hidden string lengths come last ----
void cfd2lcs_init_(int* comm, int* n, int* offset, lcsdata_t* x, lcsdata_t* y, lcsdata_t* z, int* flag) {
  cfd2lcs_init_c((MPI_Comm)(comm ? *comm : 0), n, offset, x, y, z, flag);
}
void cfd2lcs_diagnostic_init_(int* handle, int* type, int* res, lcsdata_t* t, lcsdata_t* h, char* label,
                              size_t namelen) {
  std::string l = trim_label(label, namelen);
  *handle = cfd2lcs_diagnostic_init_c(*type, *res, *t, *h, &l[0]);
}
void cfd2lcs_update_(int* n, lcsdata_t* u, lcsdata_t* v, lcsdata_t* w, lcsdata_t* time) {
  cfd2lcs_update_c(n, u, v, w, *time);
}
void cfd2lcs_diagnostic_destroy_(int* handle) { cfd2lcs_diagnostic_destroy_c(*handle); }
void cfd2lcs_set_option_(char* option, int* val, size_t namelen) {
  std::string o = trim_label(option, namelen);
  cfd2lcs_set_option_c(&o[0], *val);
}
void cfd2lcs_set_param_(char* param, lcsdata_t* val, size_t namelen) {
  std::string p = trim_label(param, namelen);
  cfd2lcs_set_param_c(&p[0], *val);
}
void cfd2lcs_finalize_(void) { cfd2lcs_finalize_c(); }

}  // extern "C"
