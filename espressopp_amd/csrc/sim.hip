// Host runtime of the MI355X short-range MD step and its C ABI (include/espp_amd.h).
//
// One espp_sim owns the device state of one periodic system and drives the
// kernels of kernels.hip on one HIP stream.  The control flow restates
// VelocityVerlet::run (src/integrator/VelocityVerlet.cpp:63-169): a full force
// evaluation with the Langevin heat-up at run() start, then per step integrate1,
// the skin/2 rebuild trigger on the accumulated maximum displacement, the force
// evaluation in interaction order, the thermostat and integrate2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/espp_amd.h"
#include "espp_dev.h"
#include "host_util.h"
#include "kernels.h"
#include "mlp.h"

using namespace espp;
using namespace espp_host;

thread_local std::string espp_host::g_err;

namespace {

constexpr size_t kPosqPad = 64;   // spare fixed-point mirror entries after the last particle

// Scale of the list build's 16-bit copies for the grid (q16_scale_cubic; 0 for
// a grid that is not cubic).
double q16_scale(const Geom& g, double disp, double top) {
  if (!(g.L[0] == g.L[1] && g.L[1] == g.L[2] && g.n[0] == g.n[1] && g.n[1] == g.n[2])) return 0.0;
  return q16_scale_cubic(g.L[0], g.n[0], disp, top);
}

// Device scratch of espp_ia_energy_noalloc: per-workgroup partial sums of an
// interaction's energy and virial, and their reduction in the last 8 entries.
// One copy per GPU, held until the process ends; every call on that GPU shares it.
constexpr size_t kScratchDoubles = (size_t)1 << 18;          // 2 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of this GPU's scratch, looked up once per device with that
// device current (the array's host-side name is not a device address).
double* scratch_for_device(int device) {
  if (device < 0 || device >= kMaxDevices) fail(ESPP_ERR_DEVICE, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[device] = static_cast<double*>(ptr);
  }
  return g_scratch_ptr[device];
}

struct VerletListState {
  double cut = 0.0;
  int cap = 0;          // entries per particle row (untiled) or 2 x words per row (tiled)
  bool tiled = false;   // entries are 16-bit offsets into the tile's staged region
  DBuf<int> nbr, cnt;   // indexed by particle slot
  DBuf<int> lane_slot;  // tiled: the slot each pair-kernel lane works on (lanes ordered by list length)
  int64_t builds = 0;
  int maxcount = 0;
  bool stale = true;   // slot order changed since the last build
  std::vector<std::pair<int32_t, int32_t>> excl;
  DBuf<int> ex_start, ex_ids;
  int ex_nid = -1;
  bool has_excl = false;
  bool untiled_only = false;   // a learned pair force reads this list: keep the untiled layout
};

struct FixedPairListState {
  std::vector<int32_t> a, b;
  DBuf<int> bstart, bpartner;
  int built_nid = -1;
  bool dirty = true;
};

// Rows (r, e, f) of one interaction.Tabulated potential and its cutoff.
struct HostTable {
  int itype = 0;
  std::vector<double> r, e, f;
  double cutoff = 0.0;
};

struct InteractionState {
  int kind = 0;  // 0: VerletList LJ, 1: FixedPairList FENE, 2: VerletList Tabulated, 3: FixedPairList Tabulated,
                 // 4: VerletList learned pair force (MLP)
  int list = 0;
  std::vector<LJParam> tab = std::vector<LJParam>(kMaxTypes * kMaxTypes);
  int maxset = -1;       // largest type with a potential
  DBuf<LJParam> dtab;
  int dtab_nt = 0;
  bool dirty = true;
  FENEParam fp{};
  // Tabulated: tables per type pair (kind 2) or tables[0] (kind 3)
  std::vector<std::shared_ptr<HostTable>> tables = std::vector<std::shared_ptr<HostTable>>(kMaxTypes * kMaxTypes);
  DBuf<TabParam> ttab;
  DBuf<double> tpool;
  // learned pair force: FP32 parameters, their bfloat16 operand matrices, the
  // gradient and Adam moments, per-slot densities and forward forces
  mlp::Spec spec{};
  DBuf<float> theta, grad, m1, m2, fpred;
  DBuf<unsigned short> wpack;
  DBuf<float2> rho;
  DBuf<double> loss, fref;
  bool params_set = false;
  int64_t adam_steps = 0;
  // tiled mixed LJ: its list pruned to the pairs within the largest cutoff plus
  // the prune shell at one step's positions, walked in place of the full list
  // while no pair can have come within the cutoff since (in_build: the list
  // build it was taken from, in_dist: maxDist at the time)
  DBuf<int> in_nbr, in_cnt, in_lane;
  int64_t in_build = -1;
  double in_dist = 0.0;
};

// ------------------------------------------------------------------ tables
// Per-interval coefficients of interaction.Tabulated's interpolation tables,
// computed once on the host when a potential is set.

// InterpolationLinear::spline (src/interaction/InterpolationLinear.cpp:139-150)
void table_linear(const std::vector<double>& x, const std::vector<double>& y, std::vector<double>& a,
                  std::vector<double>& b) {
  const int N = (int)x.size();
  a.assign((size_t)N, 0.0);
  b.assign((size_t)N, 0.0);
  for (int i = 0; i < N - 1; i++) {
    a[(size_t)i] = (y[(size_t)i + 1] - y[(size_t)i]) / (x[(size_t)i + 1] - x[(size_t)i]);
    b[(size_t)i] = y[(size_t)i] - a[(size_t)i] * x[(size_t)i];
  }
}

// Clamped cubic spline second derivatives, tridiagonal sweep of
// InterpolationCubic::spline (src/interaction/InterpolationCubic.cpp:170-215);
// an end slope above 0.99e30 selects the natural end.
void table_cubic(const std::vector<double>& x, const std::vector<double>& y, double yp1, double ypn,
                 std::vector<double>& y2) {
  const int n = (int)x.size();
  std::vector<double> u((size_t)n, 0.0);
  y2.assign((size_t)n, 0.0);
  if (yp1 > 0.99e30) {
    y2[0] = 0.0;
    u[0] = 0.0;
  } else {
    y2[0] = -0.5;
    u[0] = (3.0 / (x[1] - x[0])) * ((y[1] - y[0]) / (x[1] - x[0]) - yp1);
  }
  for (int i = 1; i < n - 1; i++) {
    const size_t k = (size_t)i;
    const double sig = (x[k] - x[k - 1]) / (x[k + 1] - x[k - 1]);
    const double p = sig * y2[k - 1] + 2.0;
    y2[k] = (sig - 1.0) / p;
    u[k] = (y[k + 1] - y[k]) / (x[k + 1] - x[k]) - (y[k] - y[k - 1]) / (x[k] - x[k - 1]);
    u[k] = (6.0 * u[k] / (x[k + 1] - x[k - 1]) - sig * u[k - 1]) / p;
  }
  double qn, un;
  const size_t m = (size_t)n;
  if (ypn > 0.99e30) {
    qn = 0.0;
    un = 0.0;
  } else {
    qn = 0.5;
    un = (3.0 / (x[m - 1] - x[m - 2])) * (ypn - (y[m - 1] - y[m - 2]) / (x[m - 1] - x[m - 2]));
  }
  y2[m - 1] = (un - qn * u[m - 2]) / (qn * y2[m - 2] + 1.0);
  for (int k = n - 2; k >= 0; k--) y2[(size_t)k] = y2[(size_t)k] * y2[(size_t)k + 1] + u[(size_t)k];
}

// InterpolationAkima::getSlope (src/interaction/InterpolationAkima.hpp:108-120)
double akima_slope(double m1, double m2, double m3, double m4) {
  if ((m1 == m2) && (m3 == m4)) return (m2 + m3) / 2.0;
  return (std::fabs(m4 - m3) * m2 + std::fabs(m2 - m1) * m3) / (std::fabs(m4 - m3) + std::fabs(m2 - m1));
}

// Akima coefficients with two parabola-extrapolated points on each side
// (InterpolationAkima::spline, src/interaction/InterpolationAkima.cpp:136-214); N >= 4.
void table_akima(const std::vector<double>& x, const std::vector<double>& y, std::vector<double>& p0,
                 std::vector<double>& p1, std::vector<double>& p2, std::vector<double>& p3) {
  const size_t N = x.size();
  std::vector<double> t(N, 0.0);
  double temp = (x[1] - x[0]) / (x[2] - x[0]);
  temp = temp * temp;
  double g0 = y[0];
  double g1 = ((y[1] - y[0]) - temp * (y[2] - y[0])) / ((x[1] - x[0]) - temp * (x[2] - x[0]));
  double g2 = ((y[2] - y[0]) - g1 * (x[2] - x[0])) / ((x[2] - x[0]) * (x[2] - x[0]));
  const double x1 = x[0] - (x[2] - x[0]);
  const double x2 = x[1] - (x[2] - x[0]);
  const double y1 = g0 + g1 * (x1 - x[0]) + g2 * (x1 - x[0]) * (x1 - x[0]);
  const double y2 = g0 + g1 * (x2 - x[0]) + g2 * (x2 - x[0]) * (x2 - x[0]);
  double m1 = (y2 - y1) / (x2 - x1);
  double m2 = (y[0] - y2) / (x[0] - x2);
  double m3 = (y[1] - y[0]) / (x[1] - x[0]);
  double m4 = (y[2] - y[1]) / (x[2] - x[1]);
  t[0] = akima_slope(m1, m2, m3, m4);
  double m5 = (y[3] - y[2]) / (x[3] - x[2]);
  t[1] = akima_slope(m2, m3, m4, m5);
  temp = (x[N - 2] - x[N - 1]) / (x[N - 3] - x[N - 1]);
  temp = temp * temp;
  g0 = y[N - 1];
  g1 = ((y[N - 2] - y[N - 1]) - temp * (y[N - 3] - y[N - 1])) / ((x[N - 2] - x[N - 1]) - temp * (x[N - 3] - x[N - 1]));
  g2 = ((y[N - 3] - y[N - 1]) - g1 * (x[N - 3] - x[N - 1])) / ((x[N - 3] - x[N - 1]) * (x[N - 3] - x[N - 1]));
  const double x4 = x[N - 2] + (x[N - 1] - x[N - 3]);
  const double x5 = x[N - 1] + (x[N - 1] - x[N - 3]);
  const double y4 = g0 + g1 * (x4 - x[N - 1]) + g2 * (x4 - x[N - 1]) * (x4 - x[N - 1]);
  const double y5 = g0 + g1 * (x5 - x[N - 1]) + g2 * (x5 - x[N - 1]) * (x5 - x[N - 1]);
  m1 = (y[N - 3] - y[N - 4]) / (x[N - 3] - x[N - 4]);
  m2 = (y[N - 2] - y[N - 3]) / (x[N - 2] - x[N - 3]);
  m3 = (y[N - 1] - y[N - 2]) / (x[N - 1] - x[N - 2]);
  m4 = (y4 - y[N - 1]) / (x4 - x[N - 1]);
  m5 = (y5 - y4) / (x5 - x4);
  t[N - 2] = akima_slope(m1, m2, m3, m4);
  t[N - 1] = akima_slope(m2, m3, m4, m5);
  for (size_t i = 2; i + 2 < N; i++) {
    m1 = (y[i - 1] - y[i - 2]) / (x[i - 1] - x[i - 2]);
    m2 = (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
    m3 = (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
    m4 = (y[i + 2] - y[i + 1]) / (x[i + 2] - x[i + 1]);
    t[i] = akima_slope(m1, m2, m3, m4);
  }
  p0.assign(N, 0.0); p1.assign(N, 0.0); p2.assign(N, 0.0); p3.assign(N, 0.0);
  for (size_t i = 0; i + 1 < N; i++) {
    const double h = x[i + 1] - x[i];
    p0[i] = y[i];
    p1[i] = t[i];
    p2[i] = (3.0 * (y[i + 1] - y[i]) / h - 2.0 * t[i] - t[i + 1]) / h;
    p3[i] = (t[i] + t[i + 1] - 2.0 * (y[i + 1] - y[i]) / h) / (h * h);
  }
}

// Pool rows {r_i, c0, c1, c2, c3} of one table: N force rows, then N energy rows.
void table_rows(const HostTable& t, std::vector<double>& pool) {
  const size_t N = t.r.size();
  auto emit = [&](const std::vector<double>& y, const std::vector<double>* dy) {
    std::vector<double> c0, c1, c2, c3;
    if (t.itype == 1) {
      table_linear(t.r, y, c0, c1);
    } else if (t.itype == 2) {
      table_akima(t.r, y, c0, c1, c2, c3);
    } else {
      // energy: end slopes are minus the tabulated forces; force: secant end slopes
      std::vector<double> y2;
      if (dy) table_cubic(t.r, y, -(*dy)[0], -(*dy)[N - 1], y2);
      else table_cubic(t.r, y, (y[1] - y[0]) / (t.r[1] - t.r[0]), (y[N - 1] - y[N - 2]) / (t.r[N - 1] - t.r[N - 2]), y2);
      c0 = y;
      c1 = y2;
    }
    for (size_t i = 0; i < N; ++i) {
      pool.push_back(t.r[i]);
      if (t.itype == 3) {
        const size_t j = std::min(i + 1, N - 1);
        pool.push_back(c0[i]); pool.push_back(c1[i]); pool.push_back(c0[j]); pool.push_back(c1[j]);
      } else {
        pool.push_back(c0[i]); pool.push_back(c1[i]);
        pool.push_back(c2.empty() ? 0.0 : c2[i]); pool.push_back(c3.empty() ? 0.0 : c3[i]);
      }
    }
  };
  emit(t.f, nullptr);
  emit(t.e, &t.f);
}

// host -> device copy ordered on the handle's (non-blocking) stream; returns when done
void upload_bytes(void* dst, const void* src, size_t bytes, hipStream_t stream, const char* what) {
  if (!bytes) return;
  hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream), what);
  hip_check(hipStreamSynchronize(stream), what);
}

}  // namespace

struct espp_sim {
  Geom g{};
  double skin = 0.0;
  uint64_t seed = 0;
  int precision = ESPP_FP64;
  int device = 0;
  hipStream_t stream = nullptr;

  int64_t n = 0;
  int maxtype = 0;
  DBuf<double4> pos, vel, pos2, vel2;
  DBuf<double> fx, fy, fz, fx2, fy2, fz2;
  DBuf<int> id, id2;
  DBuf<int4> img, img2;
  DBuf<uint4> posq;
  DBuf<uint4> posq2;              // the next step's mirror while a fused force kernel reads posq
  DBuf<float4> posf;              // FP32 positions, candidates of the tiled list build
  DBuf<uint2> posh;               // 16-bit modular positions (q16_pack), the same for cubic grids
  DBuf<int> cell_of, rank, order, cell_of_slot, cell_count, cstart;
  DBuf<int> xbin;                 // per cell, the x sub-bin starts of the tiled list build
  DBuf<int> scan_sum, scan_off;   // per-chunk totals and offsets of the cell scan
  DBuf<double> blockmax;          // per-block max |dx|^2 of the fused step
  DBuf<int> slot_of_id;
  DBuf<int> hcell;                // per slot: cell of the caller's grid (halfCellInt > 1)
  int nid = 0;

  // tile plan of the current slot order (tiled list build and force kernels)
  int lds_cap = 65536;          // dynamic+static LDS bytes one workgroup may use
  int T = 0;                    // particles per tile; 0 when no plan fits
  bool plan_valid = false;
  bool plan_pending = false;    // a tentative plan was launched; its result is not read yet
  int plan_tried = 0;           // tile size of the pending tentative plan
  int plan_failed = 0;          // tile size whose plan failed for the current slot order
  int plan_hint = 0;            // tile size of the last plan that fitted
  int large_fail_streak = 0;    // tentative plans of the largest size that failed in a row
  int64_t plan_count = 0;       // tentative plans launched
  bool nan_pending = false;     // the binning's lost-particle flag is not read yet
  int max_region = 0;           // largest staged region (particles)
  DBuf<TileRun> runs;
  DBuf<int> nruns;
  std::vector<int32_t> host_ids;  // ids present (unsorted)
  bool sorted = false;

  DBuf<Ctl> ctl;
  Ctl* hctl = nullptr;
  // per-step max displacement, written by the device into coherent pinned memory
  unsigned long long* hmax = nullptr;
  unsigned long long* dmax = nullptr;     // device view of hmax
  DBuf<double> partial, red;
  double* hred = nullptr;
  // windows of steps launched without a host wait (device-side skin check)
  DBuf<SpecState> spec;
  SpecState* hspec = nullptr;
  int spec_max = 32;               // longest window (0: wait on the host after every step)
  double spec_rate = 0.0;          // recent max displacement per step
  int64_t spec_windows = 0, spec_launched = 0, spec_ran = 0;
  double* scratch_dev = nullptr;   // this device's scratch (espp_ia_energy_noalloc)

  std::vector<VerletListState*> vls;
  std::vector<FixedPairListState*> fpls;
  std::vector<InteractionState*> ias;
  std::vector<int> active;   // System::shortRangeInteractions, in addInteraction order

  int64_t step = 0;
  double maxDist = 0.0;
  // width of the shell kept by the pruned LJ lists (0: no pruning, the default;
  // below 0: skin / 3)
  double prune_shell = 0.0;
  int64_t prunes = 0, pruned_walks = 0;   // step kernels that pruned / walked a pruned list
  double disp_since_bin = 0.0;    // bound on any particle's displacement since the last binning
  double coord_bound = 0.0;       // bound on |coordinate| of the current positions
  bool bin_folded = false;        // the last binning folded every position into the box
  bool resortFlag = true;
  int64_t nResorts = 0;

  bool lang = false;
  double gamma = 0.0, temperature = 0.0;
  std::vector<int32_t> lang_excl;
  DBuf<unsigned char> excl_flag;
  int excl_nid = -1;
  bool excl_dirty = true;
  uint64_t therm_calls = 0;

  double timers[10] = {0};

  // optional per-launch timing with HIP events on this handle's stream
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_open;
  int64_t prof_count[8] = {0};
  double prof_ms[8] = {0};

  hipEvent_t take_event() {
    if (!ev_pool.empty()) {
      hipEvent_t e = ev_pool.back();
      ev_pool.pop_back();
      return e;
    }
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    return e;
  }
  hipEvent_t prof_begin() {
    if (!prof) return nullptr;
    hipEvent_t a = take_event();
    hip_check(hipEventRecord(a, stream), "hipEventRecord");
    return a;
  }
  void prof_end(int kernel, hipEvent_t a) {
    if (!prof || !a) return;
    hipEvent_t b = take_event();
    hip_check(hipEventRecord(b, stream), "hipEventRecord");
    ev_open.push_back({kernel, {a, b}});
  }
  void prof_collect() {
    if (ev_open.empty()) return;
    sync();
    for (auto& r : ev_open) {
      float ms = 0.f;
      hip_check(hipEventElapsedTime(&ms, r.second.first, r.second.second), "hipEventElapsedTime");
      prof_count[r.first]++;
      prof_ms[r.first] += ms;
      ev_pool.push_back(r.second.first);
      ev_pool.push_back(r.second.second);
    }
    ev_open.clear();
  }

  ~espp_sim() {
    for (auto& r : ev_open) { (void)hipEventDestroy(r.second.first); (void)hipEventDestroy(r.second.second); }
    for (auto e : ev_pool) (void)hipEventDestroy(e);
    for (auto* v : vls) delete v;
    for (auto* f : fpls) delete f;
    for (auto* a : ias) delete a;
    if (hctl) (void)hipHostFree(hctl);
    if (hmax) (void)hipHostFree(hmax);
    if (hred) (void)hipHostFree(hred);
    if (hspec) (void)hipHostFree(hspec);
    if (stream) (void)hipStreamDestroy(stream);
  }

  bool mixed() const { return precision == ESPP_MIXED; }

  void sync() { hip_check(hipStreamSynchronize(stream), "stream synchronize"); }

  // Wait until the device has written *hmax (it was set to kNoValue before the
  // launch).  The stream is queried now and then so that a failed launch or an
  // idle stream without a value ends the wait with an error.
  static constexpr unsigned long long kNoValue = ~0ull;   // never the bits of a max |dx|^2 >= 0
  unsigned long long wait_host_value() {
    for (unsigned long it = 1;; ++it) {
      const unsigned long long v = __atomic_load_n(hmax, __ATOMIC_ACQUIRE);
      if (v != kNoValue) return v;
      if ((it & 1023) == 0) {
        const hipError_t e = hipStreamQuery(stream);
        if (e == hipSuccess) {
          const unsigned long long w = __atomic_load_n(hmax, __ATOMIC_ACQUIRE);
          if (w != kNoValue) return w;
          fail(ESPP_ERR_DEVICE, "the integrate kernel finished without reporting its displacement");
        }
        if (e != hipErrorNotReady) hip_check(e, "integrate");
      }
      __builtin_ia32_pause();
    }
  }

  void read_ctl() {
    hip_check(hipMemcpyAsync(hctl, ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, stream), "ctl readback");
    sync();
  }

  // ------------------------------------------------------------ particles
  void reserve_particles(int64_t need) {
    if ((size_t)need <= pos.n && pos.p) return;
    size_t m = std::max<size_t>((size_t)need, pos.n ? pos.n * 2 : 1024);
    size_t keep = (size_t)n;
    pos.grow_keep(m, keep, stream); vel.grow_keep(m, keep, stream);
    fx.grow_keep(m, keep, stream); fy.grow_keep(m, keep, stream); fz.grow_keep(m, keep, stream);
    id.grow_keep(m, keep, stream); img.grow_keep(m, keep, stream);
    if (g.hci > 1) {
      hcell.grow_keep(m, keep, stream);
      g.hcell = hcell.p;
    }
    pos2.resize(m); vel2.resize(m); fx2.resize(m); fy2.resize(m); fz2.resize(m);
    id2.resize(m); img2.resize(m);
    cell_of.resize(m); rank.resize(m); order.resize(m); cell_of_slot.grow_keep(m, keep, stream);
    g.cslot = cell_of_slot.p;
    // the tiled list build reads up to 31 entries past a candidate range
    posq.grow_keep(m + kPosqPad, keep, stream);
    posq2.resize(m + kPosqPad);
    posf.resize(m + kPosqPad);
    if (g.q16inv > 0.0) posh.resize(m + kPosqPad);
  }

  int64_t ncell() const { return (int64_t)g.n[0] * g.n[1] * g.n[2]; }

  void ensure_id_map() {
    int need = nid > 0 ? nid : 1;
    if (slot_of_id.n < (size_t)need) {
      slot_of_id.resize((size_t)need);
    }
  }

  // Fold (optionally), bin into cells, sort each cell by id and permute every
  // per-particle array into cell order.
  // With defer_check the lost-particle flag is read with the list build's
  // control block (one read per rebuild) instead of here.
  void bin(bool do_fold, bool carry_forces = true, bool defer_check = false) {
    int ni = (int)n;
    hipEvent_t pe = prof_begin();
    launch_reset_ctl(stream, ctl.p);
    launch_fill_int(stream, (int)ncell(), cell_count.p, 0);
    ensure_id_map();
    launch_fill_int(stream, nid, slot_of_id.p, -1);
    if (ni > 0) {
      Geom gg = g;
      launch_fold_cell(stream, ni, pos.p, img.p, cell_of.p, rank.p, cell_count.p, gg, ctl.p, do_fold ? 1 : 0);
    }
    launch_scan_multi(stream, (int)ncell(), cell_count.p, cstart.p, scan_sum.p, scan_off.p);
    if (ni > 0) {
      launch_scatter(stream, ni, cell_of.p, rank.p, cstart.p, order.p);
      launch_sort_cells(stream, (int)ncell(), cstart.p, order.p, pos.p, ctl.p);
      launch_permute(stream, ni, order.p, cell_of.p, pos.p, vel.p, fx.p, fy.p, fz.p, id.p, img.p, pos2.p,
                     vel2.p, carry_forces ? fx2.p : nullptr, fy2.p, fz2.p, id2.p, img2.p, cell_of_slot.p,
                     slot_of_id.p, posq.p, posf.p, posh.p, g);
      pos.swap(pos2); vel.swap(vel2);
      if (carry_forces) { fx.swap(fx2); fy.swap(fy2); fz.swap(fz2); }
      id.swap(id2); img.swap(img2);
      launch_xbins(stream, ni, (int)ncell(), cstart.p, cell_of_slot.p, posf.p, g, xbin.p);
    }
    prof_end(4, pe);
    disp_since_bin = 0.0;
    plan_valid = false;
    for (auto* v : vls) v->stale = true;
    if (do_fold) coord_bound = std::max(g.L[0], std::max(g.L[1], g.L[2]));
    bin_folded = do_fold;
    sorted = true;
    if (defer_check) {
      nan_pending = true;
    } else {
      read_ctl();
      check_nanpos();
    }
  }

  void check_nanpos() {
    nan_pending = false;
    if (hctl->nanpos)
      fail(ESPP_ERR_NAN, "particle has moved to outer space (one or more coordinates are nan or infinite)");
  }

  void ensure_sorted() {
    if (!sorted) bin(false);
  }

  // entries of a tile's staged region the plan allows (the force kernels' LDS)
  int plan_budget() const { return (lds_cap - 12288) / (mixed() ? 16 : 32); }

  // Tile plan for the current slot order: the largest tile size whose staged
  // regions fit the LDS budget; T = 0 selects the untiled kernels.
  // With tentative, one tile size is launched without waiting for its result:
  // the list build that follows (which returns at once when the plan failed)
  // reads plan_fail and max_region with its own control block, and only when
  // the plan failed are the other sizes searched, largest first.  The tentative
  // size is the largest, or, after the largest failed twice in a row, the last
  // size that fitted (the largest is tried again every 16th plan).
  void ensure_plan(bool tentative = false) {
    if (plan_valid && plan_pending && !tentative) {
      // read the tentative plan's result now
      read_ctl();
      if (nan_pending) check_nanpos();
      if (tentative_plan_done(!hctl->plan_fail)) return;
    }
    if (plan_valid) return;
    T = 0;
    max_region = 0;
    plan_valid = true;
    plan_pending = false;
    if (n == 0) return;
    // the tiled list build assumes at least 3 cells per dimension
    if (g.n[0] < 3 || g.n[1] < 3 || g.n[2] < 3) return;
    const int budget = plan_budget();
    const int sizes_mixed[3] = {256, 128, 64}, sizes_fp64[2] = {128, 64};
    const int* sizes = mixed() ? sizes_mixed : sizes_fp64;
    const int nsz = mixed() ? 3 : 2;
    auto launch_plan = [&](int tt) {
      const int ntiles = (int)((n + tt - 1) / tt);
      runs.resize((size_t)ntiles * kMaxRuns);
      nruns.resize((size_t)ntiles);
      launch_reset_ctl(stream, ctl.p, kCtlPlan);
      launch_tile_plan(stream, (int)n, tt, ntiles, cell_of_slot.p, cstart.p, g, budget, runs.p, nruns.p, ctl.p);
    };
    if (tentative) {
      int tt = sizes[0];
      if (large_fail_streak >= 2 && plan_hint > 0 && plan_count % 16 != 0) tt = plan_hint;
      ++plan_count;
      launch_plan(tt);
      T = tt;
      plan_tried = tt;
      plan_pending = true;
      return;
    }
    const int skip = plan_failed;   // a size whose plan just failed for this slot order
    plan_failed = 0;
    for (int q = 0; q < nsz; ++q) {
      const int tt = sizes[q];
      if (tt == skip) continue;
      launch_plan(tt);
      read_ctl();
      if (nan_pending) check_nanpos();
      if (!hctl->plan_fail) {
        T = tt;
        max_region = hctl->max_region;
        plan_hint = tt;
        return;
      }
    }
  }

  // Bookkeeping once a tentative plan's result is read.  Returns true when it
  // fitted (T and max_region are then final); otherwise the plan is invalid and
  // the next ensure_plan() searches the other sizes.
  bool tentative_plan_done(bool fitted) {
    plan_pending = false;
    const int largest = mixed() ? 256 : 128;
    if (plan_tried == largest) large_fail_streak = fitted ? 0 : large_fail_streak + 1;
    if (fitted) {
      max_region = hctl->max_region;
      plan_hint = T;
      return true;
    }
    plan_valid = false;
    plan_failed = plan_tried;
    return false;
  }

  // neighbour lists hold slot indices: rebuild every list the last binning invalidated
  void ensure_lists() {
    ensure_sorted();
    for (auto* v : vls)
      if (v->stale) build_vl(*v);
  }

  // ------------------------------------------------------------ lists
  void build_vl(VerletListState& v) {
    const double cutVerlet = v.cut + skin;
    const double cutsq = cutVerlet * cutVerlet;
    for (int d = 0; d < 3; ++d) {
      double cs = g.L[d] / g.n[d];
      if (cs < cutVerlet)
        fail(ESPP_ERR_ARG, "cell size " + std::to_string(cs) + " in direction " + std::to_string(d) +
                               " is smaller than cutoff+skin " + std::to_string(cutVerlet));
    }
    if (v.has_excl && v.ex_nid != nid) upload_exclusions(v);
    ensure_plan(true);
    const int ngroups = (int)((n + 63) / 64);
    if (v.cap == 0) {
      double vol = g.L[0] * g.L[1] * g.L[2];
      double nbar = 4.0 / 3.0 * M_PI * cutVerlet * cutVerlet * cutVerlet * (double)n / vol;
      int cap = (int)std::ceil(1.4 * nbar + 24.0);
      cap = std::min<int64_t>(cap, std::max<int64_t>(n, 8));
      v.cap = (cap + 7) & ~7;
    }
    v.cnt.resize((size_t)std::max<int64_t>(n, 1));
    v.tiled = T > 0 && !v.untiled_only;
    // FP32 prefilter window of the tiled build.  Its coordinates (a position, or
    // one shifted by a box length) are below M in magnitude; each is rounded to
    // FP32 once (half an ulp of M), so a difference is off by at most one ulp and
    // r^2 by at most 2 sqrt(3) (rc+skin) ulp, plus a few 1e-7 of FP32 rounding of
    // r^2 itself.  The window takes twice that.
    const double lmax = std::max(g.L[0], std::max(g.L[1], g.L[2]));
    const double M = coord_bound + lmax + cutVerlet;
    const double ulp = std::ldexp(1.0, std::ilogb(M) - 23);
    const double win = 1e-6 + 4.0 * std::sqrt(3.0) * ulp / std::max(cutVerlet, 1e-30);
    const float cut_lo_q = (float)(cutsq * (1.0 - win)), cut_hi_q = (float)(cutsq * (1.0 + win));
    // particles moved since the binning sorted the cells by x: widen the x clip
    const float clip_margin = disp_since_bin > 0.0 ? (float)(2.0 * disp_since_bin * (1.0 + 1e-6) + 4.0 * ulp) : 0.0f;
    // rows clipped in y and z when the binning folded every position into the box
    // (each cell then holds only particles of its own range); the margin covers the
    // particles' moves since, negative turns the clip off
    const float row_margin = bin_folded ? (float)(disp_since_bin * (1.0 + 1e-6) + 4.0 * ulp) : -1.0f;
    // 16-bit candidate copies (cubic grids, positions folded at the binning) with
    // a scale for the drift since the binning, or FP32 ones
    const double q16 = (v.tiled && bin_folded) ? q16_scale(g, disp_since_bin, std::sqrt((double)cut_hi_q) * 1.0001)
                                               : 0.0;
    // the candidate copies are the binning's until a particle moves
    if (v.tiled && (disp_since_bin > 0.0 || (q16 > 0.0 && q16 != g.q16inv)))
      launch_posf(stream, (int)n, pos.p, posf.p, posh.p, q16);
    int q16_tlo = 0, q16_w = 0;
    if (q16 > 0.0) q16_thresholds(cutVerlet, q16, q16_tlo, q16_w);
    for (int attempt = 0; attempt < 4; ++attempt) {
      launch_reset_ctl(stream, ctl.p, kCtlBuild);
      hipEvent_t pe = prof_begin();
      if (v.tiled) {
        // 16-bit entries, eight per 16-byte group: cap / 2 ints per particle
        v.nbr.resize((size_t)std::max(ngroups, 1) * (size_t)(v.cap / 2) * 64);
        v.lane_slot.resize((size_t)std::max<int64_t>(n, 1));
        launch_nbr_build_tiled(stream, (int)n, T, posf.p, pos.p, id.p, cell_of_slot.p, cstart.p, g,
                               cutsq, cut_lo_q, cut_hi_q, runs.p, nruns.p, v.has_excl ? v.ex_start.p : nullptr,
                               v.has_excl ? v.ex_ids.p : nullptr, (uint4*)v.nbr.p, v.cnt.p, v.cap / 8,
                               v.lane_slot.p, clip_margin, row_margin, xbin.p, posh.p, q16, q16_tlo, q16_w,
                               ctl.p);
      } else {
        v.nbr.resize((size_t)std::max(ngroups, 1) * (size_t)v.cap * 64);
        launch_nbr_build(stream, (int)n, pos.p, id.p, cell_of_slot.p, cstart.p, g, cutsq,
                         v.has_excl ? v.ex_start.p : nullptr, v.has_excl ? v.ex_ids.p : nullptr, v.nbr.p,
                         v.cnt.p, v.cap, ctl.p);
      }
      prof_end(1, pe);
      // one control-block read per rebuild: the binning's lost-particle flag,
      // the tentative tile plan and this build's capacity
      read_ctl();
      if (nan_pending) check_nanpos();
      if (plan_pending && !tentative_plan_done(!hctl->plan_fail)) {
        // the tentative tile size did not fit: search the others, then build again
        ensure_plan();
        v.tiled = T > 0 && !v.untiled_only;
        --attempt;
        continue;
      }
      v.maxcount = hctl->maxcount;
      if (!hctl->overflow) break;
      v.cap = ((int)std::ceil(v.maxcount * 1.2) + 8 + 7) & ~7;
    }
    v.stale = false;
    v.builds++;
  }

  void upload_exclusions(VerletListState& v) {
    sync();
    std::vector<int> start((size_t)nid + 1, 0), ids;
    std::vector<std::vector<int>> adj((size_t)nid);
    for (auto& e : v.excl) {
      if (e.first < nid && e.second < nid && e.first >= 0 && e.second >= 0) {
        adj[(size_t)e.first].push_back(e.second);
        adj[(size_t)e.second].push_back(e.first);
      }
    }
    for (int i = 0; i < nid; ++i) {
      std::sort(adj[(size_t)i].begin(), adj[(size_t)i].end());
      adj[(size_t)i].erase(std::unique(adj[(size_t)i].begin(), adj[(size_t)i].end()), adj[(size_t)i].end());
      start[(size_t)i + 1] = start[(size_t)i] + (int)adj[(size_t)i].size();
    }
    ids.reserve((size_t)start[(size_t)nid]);
    for (int i = 0; i < nid; ++i) ids.insert(ids.end(), adj[(size_t)i].begin(), adj[(size_t)i].end());
    v.ex_start.resize(start.size());
    v.ex_ids.resize(std::max<size_t>(ids.size(), 1));
    upload_bytes(v.ex_start.p, start.data(), start.size() * sizeof(int), stream, "excl upload");
    if (!ids.empty())
      upload_bytes(v.ex_ids.p, ids.data(), ids.size() * sizeof(int), stream, "excl upload");
    v.ex_nid = nid;
  }

  void upload_bonds(FixedPairListState& f) {
    sync();
    std::vector<int> deg((size_t)nid + 1, 0);
    for (size_t k = 0; k < f.a.size(); ++k) {
      if (f.a[k] < nid && f.b[k] < nid) { deg[(size_t)f.a[k] + 1]++; deg[(size_t)f.b[k] + 1]++; }
    }
    std::partial_sum(deg.begin(), deg.end(), deg.begin());
    std::vector<int> fill(deg.begin(), deg.end() - 1), part((size_t)std::max(deg[(size_t)nid], 1));
    for (size_t k = 0; k < f.a.size(); ++k) {
      if (f.a[k] < nid && f.b[k] < nid) {
        part[(size_t)fill[(size_t)f.a[k]]++] = f.b[k];
        part[(size_t)fill[(size_t)f.b[k]]++] = f.a[k];
      }
    }
    f.bstart.resize(deg.size());
    f.bpartner.resize(part.size());
    upload_bytes(f.bstart.p, deg.data(), deg.size() * sizeof(int), stream, "bond upload");
    upload_bytes(f.bpartner.p, part.data(), part.size() * sizeof(int), stream, "bond upload");
    f.built_nid = nid;
    f.dirty = false;
  }

  void decompose(bool carry_forces = true) {
    bin(true, carry_forces, !vls.empty());
    for (auto* v : vls) build_vl(*v);
  }

  // ------------------------------------------------------------ interactions
  int lj_ntypes(const InteractionState& a) const { return std::max(a.maxset, maxtype) + 1; }

  void upload_lj(InteractionState& a) {
    int nt = lj_ntypes(a);
    if (nt > kMaxTypes) fail(ESPP_ERR_ARG, "at most " + std::to_string(kMaxTypes) + " particle types are supported");
    if (!a.dirty && a.dtab_nt == nt) return;
    sync();
    std::vector<LJParam> t((size_t)nt * nt);
    // fixed-point units: r^2 unit q2 = (L / 2^(32+k))^2; ff1 ~ u^7, ff2 ~ u^4, u = 1/r^2
    const double q2 = g.qs2;
    const double qmax = std::max(g.qsd[0], std::max(g.qsd[1], g.qsd[2]));
    for (int i = 0; i < nt; ++i)
      for (int j = 0; j < nt; ++j) {
        LJParam p = a.tab[(size_t)i * kMaxTypes + j];
        // window of r^2 where fixed-point quantisation (<= sqrt(3) LSB per difference) could
        // flip the cutoff test
        const double rc = std::sqrt(p.drc2);
        const double win = std::isfinite(rc) ? 4.0 * std::sqrt(3.0) * qmax * rc + 1e-300 : 0.0;
        p.dlo = p.drc2 - win;
        p.dhi = p.drc2 + win;
        p.uff1 = p.dff1 / std::pow(q2, 7);
        p.uff2 = p.dff2 / std::pow(q2, 4);
        p.urc2 = p.drc2 / q2;
        p.ulo = p.dlo / q2;
        p.uhi = p.dhi / q2;
        p.usig2 = p.sig2 / q2;
        t[(size_t)i * nt + j] = p;
      }
    a.dtab.resize(t.size());
    upload_bytes(a.dtab.p, t.data(), t.size() * sizeof(LJParam), stream, "lj table upload");
    a.dtab_nt = nt;
    a.dirty = false;
    a.in_build = -1;     // cutoffs may have changed
  }

  // Whether the steps of a run can go in windows without a host wait: every
  // force kernel of a step must honour the window's stop flag (mixed LJ, FENE),
  // the last must finish the step (tiled mixed LJ, FENE), the pruned lists
  // (judged on the host) must be off, and profiling (which times each launch)
  // must be off.
  bool window_ok() const {
    if (spec_max < 2 || prof || !mixed() || active.empty() || n == 0) return false;
    const double w = prune_width();
    if (w > 0.0 && w < skin) return false;
    for (int k : active) {
      const int kind = ias[(size_t)k]->kind;
      if (kind != 0 && kind != 1) return false;
    }
    // the last interaction must be one whose kernel finishes the step
    const InteractionState& last = *ias[(size_t)active.back()];
    return last.kind == 1 || vls[(size_t)last.list]->tiled;
  }

  // ------------------------------------------------------------ pruned LJ lists
  double prune_width() const { return prune_shell < 0.0 ? skin / 3.0 : prune_shell; }

  // A pair left out of a pruned list was at least rc + w apart at the prune, in
  // fixed point (rc: the largest cutoff of the interaction, w: the shell).  Every
  // particle has moved by at most maxDist - in_dist since, and a fixed-point
  // distance is within sqrt(3) LSB of the true one, so the pair is still beyond
  // the cutoff window (rc + 2 sqrt(3) LSB) while 2 (maxDist - in_dist) + 8 sqrt(3)
  // LSB < w.
  bool pruned_valid(const InteractionState& a, const VerletListState& v) const {
    const double w = prune_width();
    const double qmax = std::max(g.qsd[0], std::max(g.qsd[1], g.qsd[2]));
    return w > 0.0 && w < skin && a.in_build == v.builds &&
           2.0 * (maxDist - a.in_dist) + 8.0 * std::sqrt(3.0) * qmax < w;
  }

  // How a mixed force walk of LJ interaction a over tiled list v runs: true to
  // walk the pruned list; false to walk the full list, having filled po when
  // that walk is to write a new pruned list.
  bool prune_plan(InteractionState& a, const VerletListState& v, PruneOut& po) {
    if (pruned_valid(a, v)) {
      ++pruned_walks;
      return true;
    }
    const double w = prune_width();
    if (w > 0.0 && w < skin) {
      a.in_nbr.resize(v.nbr.n);
      a.in_cnt.resize((size_t)std::max<int64_t>(n, 1));
      a.in_lane.resize((size_t)std::max<int64_t>(n, 1));
      po = PruneOut{(uint4*)a.in_nbr.p, a.in_cnt.p, a.in_lane.p, prune_r2(a)};
      a.in_build = v.builds;
      a.in_dist = maxDist;
      ++prunes;
    }
    return false;
  }

  // The pruned list's threshold in the kernel's units of r^2: (rc + w)^2, rc the
  // largest cutoff of the set type pairs.
  double prune_r2(const InteractionState& a) const {
    double rc = 0.0;
    for (int i = 0; i <= a.maxset; ++i)
      for (int j = 0; j <= a.maxset; ++j) {
        const LJParam& p = a.tab[(size_t)i * kMaxTypes + j];
        if (p.set && std::isfinite(p.drc2)) rc = std::max(rc, std::sqrt(p.drc2));
      }
    const double r = rc + prune_width();
    return cubic_box(g) ? r * r / g.qs2 : r * r;
  }

  // Device tables of a Tabulated interaction: TabParam per type pair (kind 2) or
  // one (kind 3), rows in a coefficient pool.
  void upload_tab(InteractionState& a) {
    const int nt = a.kind == 2 ? lj_ntypes(a) : 1;
    if (nt > kMaxTypes) fail(ESPP_ERR_ARG, "at most " + std::to_string(kMaxTypes) + " particle types are supported");
    if (!a.dirty && a.dtab_nt == nt) return;
    sync();
    std::vector<TabParam> prm((size_t)nt * nt);
    std::vector<double> pool;
    std::vector<std::pair<const HostTable*, int>> done;   // a table shared by (t1,t2) and (t2,t1) is stored once
    const double qmax = std::max(g.qsd[0], std::max(g.qsd[1], g.qsd[2]));
    for (int i = 0; i < nt; ++i)
      for (int j = 0; j < nt; ++j) {
        TabParam p{};
        const HostTable* t = a.tables[a.kind == 2 ? (size_t)i * kMaxTypes + j : 0].get();
        if (t) {
          int off = -1;
          for (auto& d : done) if (d.first == t) off = d.second;
          if (off < 0) {
            off = (int)pool.size();
            table_rows(*t, pool);
            done.push_back({t, off});
          }
          const int N = (int)t->r.size();
          p.itype = t->itype;
          p.N = N;
          p.off = off;
          p.set = 1;
          p.inner = t->r[0];
          const double delta = (t->r[(size_t)N - 1] - t->r[0]) / (N - 1);
          p.invdelta = 1.0 / delta;
          p.deltasq6 = delta * delta / 6.0;
          p.rc2 = t->cutoff * t->cutoff;
          p.shift = 0.0;
          const double win = std::isfinite(t->cutoff) ? 4.0 * std::sqrt(3.0) * qmax * t->cutoff + 1e-300 : 0.0;
          p.dlo = std::isfinite(t->cutoff) ? p.rc2 - win : INFINITY;
          p.dhi = std::isfinite(t->cutoff) ? p.rc2 + win : INFINITY;
        }
        prm[(size_t)i * nt + j] = p;
      }
    if (pool.empty()) pool.push_back(0.0);
    a.ttab.resize(prm.size());
    a.tpool.resize(pool.size());
    upload_bytes(a.ttab.p, prm.data(), prm.size() * sizeof(TabParam), stream, "table upload");
    upload_bytes(a.tpool.p, pool.data(), pool.size() * sizeof(double), stream, "table upload");
    a.dtab_nt = nt;
    a.dirty = false;
  }

  void run_interaction(InteractionState& a, bool energy, bool accum, double* partial_out,
                       const int* stop = nullptr) {
    if (a.kind == 4) {
      if (energy) fail(ESPP_ERR_STATE, "the learned pair force is not conservative: it has no energy or virial");
      if (!a.params_set) fail(ESPP_ERR_STATE, "the learned pair force has no parameters (setParameters)");
      VerletListState& v = *vls[(size_t)a.list];
      a.rho.resize((size_t)std::max<int64_t>(n, 1));
      hipEvent_t pe = prof_begin();
      mlp::launch_density(stream, (int)n, pos.p, v.nbr.p, v.cnt.p, v.cap, g, a.spec, a.rho.p);
      mlp::launch_forces(stream, (int)n, pos.p, v.nbr.p, v.cnt.p, v.cap, g, a.spec, a.rho.p, a.wpack.p, a.theta.p,
                         accum, fx.p, fy.p, fz.p);
      prof_end(0, pe);
      return;
    }
    if (a.kind == 2) {
      VerletListState& v = *vls[(size_t)a.list];
      upload_tab(a);
      hipEvent_t pe = energy ? nullptr : prof_begin();
      if (v.tiled)
        launch_tab_tiled(stream, mixed(), energy, accum, (int)n, T, max_region * (mixed() ? 16 : 32), posq.p, pos.p,
                         (const uint4*)v.nbr.p, v.cnt.p, v.cap / 8, g, runs.p, nruns.p, v.lane_slot.p, a.ttab.p,
                         a.tpool.p,
                         a.dtab_nt, fx.p, fy.p, fz.p, partial_out);
      else
        launch_tab(stream, energy, accum, (int)n, pos.p, v.nbr.p, v.cnt.p, v.cap, g, a.ttab.p, a.tpool.p, a.dtab_nt,
                   fx.p, fy.p, fz.p, partial_out);
      prof_end(0, pe);
      return;
    }
    if (a.kind == 3) {
      FixedPairListState& f = *fpls[(size_t)a.list];
      if (f.dirty || f.built_nid != nid) upload_bonds(f);
      upload_tab(a);
      hipEvent_t pe = energy ? nullptr : prof_begin();
      launch_bond_tab(stream, energy, accum, (int)n, pos.p, id.p, slot_of_id.p, nid, f.bstart.p, f.bpartner.p, g,
                      a.ttab.p, a.tpool.p, fx.p, fy.p, fz.p, partial_out);
      prof_end(5, pe);
      return;
    }
    if (a.kind == 0) {
      VerletListState& v = *vls[(size_t)a.list];
      upload_lj(a);
      bool onetype = a.dtab_nt == 1;
      hipEvent_t pe = energy ? nullptr : prof_begin();
      // mixed force walks take the pruned list while it covers the positions, and
      // otherwise prune the full list at these positions
      PruneOut po{};
      const bool inner = !energy && v.tiled && mixed() && prune_plan(a, v, po);
      if (v.tiled)
        launch_lj_tiled(stream, mixed(), onetype, energy, accum, (int)n, T,
                        max_region * (mixed() ? (onetype ? 12 : 16) : 32), posq.p,
                        pos.p, (const uint4*)(inner ? a.in_nbr.p : v.nbr.p), inner ? a.in_cnt.p : v.cnt.p,
                        v.cap / 8, g, runs.p, nruns.p, inner ? a.in_lane.p : v.lane_slot.p,
                        a.dtab.p, a.dtab_nt,
                        fx.p, fy.p, fz.p, partial_out, po.nbr ? &po : nullptr, stop);
      else
        launch_lj(stream, mixed(), onetype, energy, accum, (int)n, posq.p, pos.p, v.nbr.p, v.cnt.p, v.cap, g,
                  a.dtab.p, a.dtab_nt, fx.p, fy.p, fz.p, partial_out, stop);
      prof_end(0, pe);
    } else {
      FixedPairListState& f = *fpls[(size_t)a.list];
      if (f.dirty || f.built_nid != nid) upload_bonds(f);
      hipEvent_t pe = energy ? nullptr : prof_begin();
      launch_fene(stream, mixed(), energy, accum, (int)n, pos.p, posq.p, id.p, slot_of_id.p, nid, f.bstart.p,
                  f.bpartner.p, g, a.fp, fx.p, fy.p, fz.p, partial_out, stop);
      prof_end(5, pe);
    }
  }

  void calc_forces() {
    ensure_lists();
    bool first = true;
    for (int k : active) {
      run_interaction(*ias[(size_t)k], false, !first, nullptr);
      first = false;
    }
    if (first) launch_zero3(stream, (int)n, fx.p, fy.p, fz.p);
  }

  // Forces of a step whose last interaction, a tiled mixed LJ, also finishes the
  // step: thermostat, integrate2, and integrate1 of the next step into pos2 /
  // posq2, which then become pos / posq; the step's max |dx|^2 arrives in hmax.
  // Returns false, having launched nothing, when the last interaction is not
  // such a kernel.
  // With in_window, the step is one of a window launched without a host wait:
  // every kernel of it returns at once once the device-side skin check has
  // raised spec->stop, and the displacement goes to spec instead of hmax.
  bool calc_forces_step(StepArgs sa, bool in_window = false) {
    if (active.empty() || n == 0) return false;
    InteractionState& last = *ias[(size_t)active.back()];
    if (in_window) sa.stop = &spec.p->stop;
    if (last.kind == 1) return fene_step(last, sa, in_window);
    if (!mixed() || last.kind != 0) return false;
    ensure_lists();
    VerletListState& v = *vls[(size_t)last.list];
    if (!v.tiled) return false;
    for (size_t q = 0; q + 1 < active.size(); ++q)
      run_interaction(*ias[(size_t)active[q]], false, q > 0, nullptr, sa.stop);
    upload_lj(last);
    const bool onetype = last.dtab_nt == 1;
    blockmax.resize((size_t)std::max(lj_tiled_blocks((int)n, T), 1));
    sa.id = id.p;
    sa.vel = vel.p;
    sa.pos = pos.p;
    sa.pos_out = pos2.p;
    sa.posq_out = posq2.p;
    sa.blockmax = blockmax.p;
    // walk the pruned list while it covers the positions; otherwise walk the
    // full list and prune it at these positions
    PruneOut po{};
    const bool inner = prune_plan(last, v, po);
    if (!in_window) __atomic_store_n(hmax, kNoValue, __ATOMIC_RELEASE);
    hipEvent_t pe = prof_begin();
    launch_lj_tiled_step(stream, onetype, active.size() > 1, (int)n, T, max_region * (onetype ? 12 : 16), posq.p,
                         pos.p, (const uint4*)(inner ? last.in_nbr.p : v.nbr.p), inner ? last.in_cnt.p : v.cnt.p,
                         v.cap / 8, g, runs.p, nruns.p, inner ? last.in_lane.p : v.lane_slot.p,
                         last.dtab.p, last.dtab_nt, fx.p, fy.p, fz.p, sa, ctl.p, in_window ? nullptr : dmax,
                         po.nbr ? &po : nullptr, in_window ? spec.p : nullptr);
    prof_end(inner ? 7 : 6, pe);
    pos.swap(pos2);
    posq.swap(posq2);
    return true;
  }

  // The same with FENE bonds as the last interaction (either precision).
  bool fene_step(InteractionState& last, StepArgs sa, bool in_window = false) {
    ensure_lists();
    for (size_t q = 0; q + 1 < active.size(); ++q)
      run_interaction(*ias[(size_t)active[q]], false, q > 0, nullptr, sa.stop);
    FixedPairListState& f = *fpls[(size_t)last.list];
    if (f.dirty || f.built_nid != nid) upload_bonds(f);
    blockmax.resize((size_t)std::max(kinetic_blocks((int)n), 1));
    sa.id = id.p;
    sa.vel = vel.p;
    sa.pos = pos.p;
    sa.pos_out = pos2.p;
    sa.posq_out = mixed() ? posq2.p : nullptr;
    sa.blockmax = blockmax.p;
    if (!in_window) __atomic_store_n(hmax, kNoValue, __ATOMIC_RELEASE);
    hipEvent_t pe = prof_begin();
    launch_fene_step(stream, mixed(), active.size() > 1, (int)n, pos.p, posq.p, id.p, slot_of_id.p, nid, f.bstart.p,
                     f.bpartner.p, g, last.fp, fx.p, fy.p, fz.p, sa, ctl.p, in_window ? nullptr : dmax,
                     in_window ? spec.p : nullptr);
    prof_end(5, pe);
    pos.swap(pos2);
    if (mixed()) posq.swap(posq2);
    return true;
  }

  void upload_langevin_excl() {
    if (!excl_dirty && excl_nid == nid) return;
    sync();
    std::vector<unsigned char> f((size_t)std::max(nid, 1), 0);
    for (int32_t e : lang_excl) if (e >= 0 && e < nid) f[(size_t)e] = 1;
    excl_flag.resize(f.size());
    upload_bytes(excl_flag.p, f.data(), f.size(), stream, "langevin exclusion upload");
    excl_nid = nid;
    excl_dirty = false;
  }

  // ------------------------------------------------------------ reductions
  void reduce(int nblk, int nv, double* out) {
    launch_reduce_partials(stream, nblk, nv, partial.p, red.p);
    hip_check(hipMemcpyAsync(hred, red.p, (size_t)nv * sizeof(double), hipMemcpyDeviceToHost, stream), "reduce readback");
    sync();
    std::memcpy(out, hred, (size_t)nv * sizeof(double));
  }

  void interaction_sums(InteractionState& a, double out8[8]) {
    ensure_lists();
    bool tiled = (a.kind == 0 || a.kind == 2) && vls[(size_t)a.list]->tiled;
    int nblk = std::max(tiled ? lj_tiled_blocks((int)n, T) : lj_blocks((int)n), 1);
    partial.resize((size_t)nblk * 8);
    if (n == 0) { for (int q = 0; q < 8; ++q) out8[q] = 0.0; return; }
    run_interaction(a, true, false, partial.p);
    reduce(nblk, 8, out8);
  }

  // ------------------------------------------------------------ integrator
  void run(int nsteps, double dt) {
    using clk = std::chrono::steady_clock;
    auto t0 = clk::now();
    double t_int1 = 0, t_resort = 0, t_force = 0, t_int2 = 0;
    nResorts = 0;
    const double skinHalf = 0.5 * skin;
    const double pref1 = -gamma;
    const double pref2 = std::sqrt(24.0 * temperature * gamma / dt);
    if (lang) upload_langevin_excl();
    if (resortFlag) {
      decompose();
      maxDist = 0.0;
      resortFlag = false;
    } else {
      ensure_lists();
    }
    // recalc1 (heatUp) -> updateForces -> recalc2 (coolDown)
    calc_forces();
    if (lang) {
      launch_thermo_integrate2(stream, (int)n, id.p, vel.p, fx.p, fy.p, fz.p, 1, pref1, pref2 * std::sqrt(3.0), seed,
                               therm_calls++, excl_flag.p, excl_nid, 0, 0.5 * dt);
    }
    auto step_args = [&](bool thermo, int do_int2) {
      StepArgs sa{};
      sa.excl = excl_flag.p;
      sa.nid = excl_nid;
      sa.pref1 = pref1;
      sa.pref2 = pref2;
      sa.dt = dt;
      sa.seed = seed;
      sa.lang = thermo ? 1 : 0;
      sa.call = thermo ? therm_calls : 0ull;
      sa.do_int2 = do_int2;
      return sa;
    };
    auto add_max_dist = [&]() {
      double maxSq = 0.0;
      if (n > 0) {
        const unsigned long long bits = wait_host_value();
        std::memcpy(&maxSq, &bits, sizeof(double));
      }
      maxDist += std::sqrt(maxSq);
      disp_since_bin += std::sqrt(maxSq);
      coord_bound += std::sqrt(maxSq);
    };
    // Each step: integrate1, the displacement check (resort), forces, and the
    // thermostat with integrate2.  Integrate1 runs with the thermostat and
    // integrate2 of the previous step (k_step_fused), or inside the previous
    // step's force kernel when that kernel can finish the step (fused_prev).
    bool fused_prev = false;
    for (int i = 0; i < nsteps; ++i) {
      auto ta = clk::now();
      if (!fused_prev) {
        const bool first = (i == 0);
        StepArgs sa = step_args(!first && lang, first ? 0 : 1);
        if (!first && lang) therm_calls++;
        blockmax.resize((size_t)std::max(kinetic_blocks((int)n), 1));
        sa.id = id.p;
        sa.vel = vel.p;
        sa.pos = pos.p;
        sa.pos_out = pos.p;
        sa.posq_out = mixed() ? posq.p : nullptr;
        sa.blockmax = blockmax.p;
        hipEvent_t pe1 = prof_begin();
        __atomic_store_n(hmax, kNoValue, __ATOMIC_RELEASE);
        launch_step_fused(stream, (int)n, fx.p, fy.p, fz.p, sa, g, ctl.p, dmax);
        prof_end(2, pe1);
        add_max_dist();
      }
      auto tb = clk::now();
      t_int1 += std::chrono::duration<double>(tb - ta).count();
      // a NaN displacement also resorts, so the binning reports the lost particle
      if (!(maxDist <= skinHalf)) resortFlag = true;
      if (resortFlag) {
        decompose(false);   // forces are recomputed below
        maxDist = 0.0;
        resortFlag = false;
        nResorts++;
      }
      auto tc = clk::now();
      t_resort += std::chrono::duration<double>(tc - tb).count();
      // Steps i, i+1, ... whose force kernel finishes them are launched as one
      // window without waiting on the host: the device adds each step's max
      // displacement to maxDist and, once it passes skin/2, stops the window's
      // remaining kernels; the host then reads how many steps ran and goes on
      // with the resort check as after any step.  The window is sized to end
      // about when the list is due, from the displacement per step so far.
      const int avail = nsteps - 1 - i;
      if (avail >= 2 && window_ok()) {
        int K = spec_max;
        if (spec_rate > 0.0) K = (int)std::min<double>(spec_max, std::floor((skinHalf - maxDist) / spec_rate) + 1.0);
        K = std::max(2, std::min(K, avail));
        launch_spec_begin(stream, spec.p, maxDist, skinHalf);
        for (int k = 0; k < K; ++k) {
          StepArgs sa = step_args(lang, 1);
          sa.call = lang ? therm_calls + (uint64_t)k : 0ull;
          calc_forces_step(sa, true);
        }
        hip_check(hipMemcpyAsync(hspec, spec.p, sizeof(SpecState), hipMemcpyDeviceToHost, stream), "window readback");
        sync();
        const int d = hspec->done;
        if (d < 1 || d > K) fail(ESPP_ERR_DEVICE, "a window of steps reported " + std::to_string(d) + " steps");
        // each launch swapped the position buffers; only d steps ran
        if ((K - d) & 1) { pos.swap(pos2); posq.swap(posq2); }
        const double grow = hspec->dmax - maxDist;
        maxDist = hspec->dmax;
        disp_since_bin += grow;
        coord_bound += grow;
        if (std::isfinite(grow) && grow > 0.0) spec_rate = grow / d;
        if (lang) therm_calls += (uint64_t)d;
        ++spec_windows;
        spec_launched += K;
        spec_ran += d;
        fused_prev = true;
        t_force += std::chrono::duration<double>(clk::now() - tc).count();
        step += d;
        i += d - 1;
        continue;
      }
      fused_prev = i + 1 < nsteps && calc_forces_step(step_args(lang, 1));
      if (fused_prev) {
        if (lang) therm_calls++;
        add_max_dist();
      } else {
        calc_forces();
      }
      t_force += std::chrono::duration<double>(clk::now() - tc).count();
      step++;
    }
    if (nsteps > 0) {
      // thermostat + integrate2 of the last step
      auto td = clk::now();
      hipEvent_t pe2 = prof_begin();
      launch_thermo_integrate2(stream, (int)n, id.p, vel.p, fx.p, fy.p, fz.p, lang ? 1 : 0, pref1, pref2, seed,
                               lang ? therm_calls++ : 0ull, excl_flag.p, excl_nid, 1, 0.5 * dt);
      prof_end(3, pe2);
      t_int2 += std::chrono::duration<double>(clk::now() - td).count();
    }
    sync();
    double t_run = std::chrono::duration<double>(clk::now() - t0).count();
    prof_collect();
    timers[0] = t_run;
    timers[1] = t_force;
    timers[2] = timers[3] = 0.0;
    timers[4] = timers[5] = 0.0;
    timers[6] = t_int1;
    timers[7] = t_int2;
    timers[8] = t_resort;
    timers[9] = t_run - (t_force + t_int1 + t_int2 + t_resort);
  }

  // ------------------------------------------------------------ host access
  std::vector<int64_t> slots_by_id(const std::vector<int>& ids_dev) const {
    std::vector<int64_t> s((size_t)n);
    std::iota(s.begin(), s.end(), 0);
    std::sort(s.begin(), s.end(), [&](int64_t a, int64_t b) { return ids_dev[(size_t)a] < ids_dev[(size_t)b]; });
    return s;
  }
};

// ------------------------------------------------------------------ C ABI
namespace {

// Fixed-point mirror: x -> low 32 bits of x * 2^(32+k) / L.  Differences of
// two mirrors are exact int32 and equal the minimum image modulo P = L / 2^k;
// k is the largest value keeping P >= 6 cell sizes, so every candidate of the
// 27 neighbour cells (|dx| < 2 cells plus drift) is represented unaliased.
// The binning grid is the caller's grid coarsened by halfCellInt h: floor(nh / h)
// cells, each at least h times the caller's cell, i.e. at least rc + skin for a
// grid from tools.decomp.cellGrid(..., h).
void set_geom(Geom& g, const double box[3], const int32_t cells[3], int hci) {
  g.hci = hci;
  g.hcell = nullptr;
  g.cslot = nullptr;
  for (int d = 0; d < 3; ++d) {
    g.L[d] = box[d];
    g.invL[d] = 1.0 / box[d];
    g.halfL[d] = 0.5 * box[d];
    g.nh[d] = cells[d];
    // CellGrid: cellSize = (myRight - myLeft) / n, invCellSize = 1 / cellSize (src/storage/CellGrid.cpp:52-56)
    g.hics[d] = 1.0 / ((box[d] - 0.0) / (double)cells[d]);
    g.n[d] = cells[d] / hci;
    double cs = (box[d] - 0.0) / (double)g.n[d];
    g.ics[d] = 1.0 / cs;
    int kbits = 0;
    while (kbits < 16 && (double)g.n[d] / (double)(1 << (kbits + 1)) >= 6.0) ++kbits;
    const double scale = std::ldexp(1.0, 32 + kbits);
    g.qinv[d] = scale / box[d];
    g.qs[d] = (float)(box[d] / scale);
    g.qsd[d] = box[d] / scale;
  }
  g.qs2 = g.qsd[0] * g.qsd[0];
  g.qsf = (float)g.qsd[0];
  // at the binning the list's cut is not known: the largest a cell allows
  g.q16inv = q16_scale(g, 0.0, g.L[0] / g.n[0]);
}

template <class T>
std::vector<T> download(const DBuf<T>& b, size_t count, hipStream_t stream) {
  // ordered after all work of the handle's (non-blocking) stream
  std::vector<T> h(count);
  if (count) {
    hip_check(hipMemcpyAsync(h.data(), b.p, count * sizeof(T), hipMemcpyDeviceToHost, stream), "download");
    hip_check(hipStreamSynchronize(stream), "download");
  }
  return h;
}

}  // namespace

extern "C" {

const char* espp_last_error(void) { return g_err.c_str(); }

const char* espp_version(void) { return "espp_amd 0.1 (gfx950)"; }

int espp_device_count(int32_t* count) {
  return guarded([&] {
    need(count != nullptr, "count is NULL");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) {
      (void)hipGetLastError();
      c = 0;
    }
    *count = c;
  });
}

int espp_create(const double box[3], const int32_t cell_grid[3], int32_t half_cell_int, double skin, uint64_t seed,
                int32_t precision, espp_sim** out) {
  return guarded([&] {
    need(out != nullptr && box != nullptr && cell_grid != nullptr, "NULL argument");
    *out = nullptr;
    need(half_cell_int >= 1 && half_cell_int <= 8, "halfCellInt must be between 1 and 8");
    for (int d = 0; d < 3; ++d) {
      need(box[d] > 0.0 && std::isfinite(box[d]), "box lengths must be positive and finite");
      need(cell_grid[d] > 0, "cell grid must be positive");
      // one cell (or fewer than 2 h half-cells) in a dimension would meet a particle's own
      // periodic images in the cell stencil: the reference's DomainDecomposition raises the
      // grid to 2 there (src/storage/DomainDecomposition.py:106-111)
      if (cell_grid[d] / half_cell_int < 2)
        fail(ESPP_ERR_ARG, "cell grid " + std::to_string(cell_grid[d]) + " in direction " + std::to_string(d) +
                               " is below 2 x halfCellInt (" + std::to_string(half_cell_int) +
                               "): the box must hold two cells of at least cutoff + skin");
    }
    need(skin >= 0.0, "skin must be non-negative");
    need(precision == ESPP_FP64 || precision == ESPP_MIXED, "precision must be ESPP_FP64 or ESPP_MIXED");
    int dc = 0;
    if (hipGetDeviceCount(&dc) != hipSuccess || dc == 0) {
      (void)hipGetLastError();
      fail(ESPP_ERR_DEVICE, "no HIP device available: the MI355X path needs a GPU");
    }
    auto* s = new espp_sim();
    try {
      set_geom(s->g, box, cell_grid, half_cell_int);
      s->skin = skin;
      s->seed = seed;
      s->precision = precision;
      hip_check(hipGetDevice(&s->device), "hipGetDevice");
      int lds = 0;
      if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, s->device) == hipSuccess && lds > 0)
        s->lds_cap = std::min(lds, 65536);
      hip_check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
      s->ctl.resize(1);
      hip_check(hipHostMalloc((void**)&s->hctl, sizeof(Ctl), hipHostMallocDefault), "hipHostMalloc");
      hip_check(hipHostMalloc((void**)&s->hmax, 64, hipHostMallocCoherent | hipHostMallocMapped), "hipHostMalloc");
      hip_check(hipHostGetDevicePointer((void**)&s->dmax, s->hmax, 0), "hipHostGetDevicePointer");
      hip_check(hipHostMalloc((void**)&s->hred, 16 * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
      hip_check(hipHostMalloc((void**)&s->hspec, sizeof(SpecState), hipHostMallocDefault), "hipHostMalloc");
      s->spec.resize(1);
      s->red.resize(16);
      s->scratch_dev = scratch_for_device(s->device);
      int64_t nc = s->ncell();
      need(nc < (int64_t)1 << 30, "cell grid too large");
      s->cell_count.resize((size_t)nc);
      s->cstart.resize((size_t)nc + 1);
      s->xbin.resize((size_t)nc * (kXBins + 1));
      s->scan_sum.resize((size_t)scan_blocks((int)nc) + 1);
      s->scan_off.resize((size_t)scan_blocks((int)nc) + 1);
      s->reserve_particles(1024);
      s->ensure_id_map();
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
  });
}

int espp_destroy(espp_sim* sim) {
  return guarded([&] {
    if (!sim) return;
    (void)hipStreamSynchronize(sim->stream);
    delete sim;
  });
}

int espp_set_skin(espp_sim* sim, double skin) {
  return guarded([&] {
    need(sim && skin >= 0.0, "invalid skin");
    sim->skin = skin;
  });
}

int espp_add_particles(espp_sim* s, int64_t n, const int32_t* id, const int32_t* type, const double* pos,
                       const double* vel, const double* mass) {
  return guarded([&] {
    need(s != nullptr, "NULL handle");
    need(n >= 0, "negative count");
    if (n == 0) return;
    need(id != nullptr && pos != nullptr, "id and pos are mandatory");
    need(s->n + n < (int64_t)1 << 31, "too many particles");
    // duplicate ids: the reference refuses the whole list (src/storage/Storage.py:281-290)
    std::vector<int32_t> newids(id, id + n);
    int maxid = s->nid - 1;
    for (int64_t k = 0; k < n; ++k) {
      need(id[k] >= 0, "particle ids must be non-negative");
      maxid = std::max(maxid, (int)id[k]);
    }
    {
      std::vector<unsigned char> seen((size_t)maxid + 1, 0);
      for (int32_t e : s->host_ids) seen[(size_t)e] = 1;
      for (int64_t k = 0; k < n; ++k) {
        if (seen[(size_t)id[k]]) fail(ESPP_ERR_ARG, "particle " + std::to_string(id[k]) + " already exists");
        seen[(size_t)id[k]] = 1;
      }
    }
    std::vector<double4> hp((size_t)n), hv((size_t)n);
    std::vector<int4> hi((size_t)n);
    std::vector<double> zero((size_t)n, 0.0);
    for (int64_t k = 0; k < n; ++k) {
      int t = type ? type[k] : 0;
      need(t >= 0, "particle types must be non-negative");
      s->maxtype = std::max(s->maxtype, t);
      double x[3] = {pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]};
      int im[3] = {0, 0, 0};
      // Storage::addParticle folds on insertion (src/storage/Storage.cpp:299)
      for (int d = 0; d < 3; ++d) {
        if (!(x[d] == x[d])) fail(ESPP_ERR_NAN, "particle position is nan");
        int tmp = (int)std::floor(x[d] * s->g.invL[d]);
        im[d] += tmp;
        x[d] -= (double)tmp * s->g.L[d];
      }
      double m = mass ? mass[k] : 1.0;
      need(m > 0.0, "particle mass must be positive");
      hp[(size_t)k] = make_double4(x[0], x[1], x[2], (double)t);
      hv[(size_t)k] = make_double4(vel ? vel[3 * k] : 0.0, vel ? vel[3 * k + 1] : 0.0, vel ? vel[3 * k + 2] : 0.0, m);
      hi[(size_t)k] = make_int4(im[0], im[1], im[2], 0);
    }
    s->reserve_particles(s->n + n);
    size_t off = (size_t)s->n;
    upload_bytes(s->pos.p + off, hp.data(), (size_t)n * sizeof(double4), s->stream, "upload");
    upload_bytes(s->vel.p + off, hv.data(), (size_t)n * sizeof(double4), s->stream, "upload");
    upload_bytes(s->img.p + off, hi.data(), (size_t)n * sizeof(int4), s->stream, "upload");
    upload_bytes(s->id.p + off, id, (size_t)n * sizeof(int32_t), s->stream, "upload");
    upload_bytes(s->fx.p + off, zero.data(), (size_t)n * sizeof(double), s->stream, "upload");
    upload_bytes(s->fy.p + off, zero.data(), (size_t)n * sizeof(double), s->stream, "upload");
    upload_bytes(s->fz.p + off, zero.data(), (size_t)n * sizeof(double), s->stream, "upload");
    s->n += n;
    s->coord_bound = std::max(s->coord_bound, std::max(s->g.L[0], std::max(s->g.L[1], s->g.L[2])));
    s->host_ids.insert(s->host_ids.end(), newids.begin(), newids.end());
    s->nid = std::max(s->nid, maxid + 1);
    s->sorted = false;
  });
}

int espp_decompose(espp_sim* s) {
  return guarded([&] {
    need(s != nullptr, "NULL handle");
    s->decompose();
  });
}

int espp_num_particles(espp_sim* s, int64_t* n) {
  return guarded([&] {
    need(s && n, "NULL argument");
    *n = s->n;
  });
}

int espp_get_ids(espp_sim* s, int32_t* out) {
  return guarded([&] {
    need(s && (out || s->n == 0), "NULL argument");
    std::vector<int32_t> ids = s->host_ids;
    std::sort(ids.begin(), ids.end());
    std::copy(ids.begin(), ids.end(), out);
  });
}

int espp_get_particles(espp_sim* s, int32_t field, double* out) {
  return guarded([&] {
    need(s && (out || s->n == 0), "NULL argument");
    size_t n = (size_t)s->n;
    if (n == 0) return;
    s->sync();
    std::vector<int> ids = download(s->id, n, s->stream);
    std::vector<int64_t> order = s->slots_by_id(ids);
    if (field == ESPP_FIELD_POS || field == ESPP_FIELD_TYPE) {
      std::vector<double4> p = download(s->pos, n, s->stream);
      for (size_t k = 0; k < n; ++k) {
        const double4& q = p[(size_t)order[k]];
        if (field == ESPP_FIELD_POS) { out[3 * k] = q.x; out[3 * k + 1] = q.y; out[3 * k + 2] = q.z; }
        else out[k] = q.w;
      }
    } else if (field == ESPP_FIELD_VEL || field == ESPP_FIELD_MASS) {
      std::vector<double4> v = download(s->vel, n, s->stream);
      for (size_t k = 0; k < n; ++k) {
        const double4& q = v[(size_t)order[k]];
        if (field == ESPP_FIELD_VEL) { out[3 * k] = q.x; out[3 * k + 1] = q.y; out[3 * k + 2] = q.z; }
        else out[k] = q.w;
      }
    } else if (field == ESPP_FIELD_FORCE) {
      std::vector<double> a = download(s->fx, n, s->stream), b = download(s->fy, n, s->stream),
                          c = download(s->fz, n, s->stream);
      for (size_t k = 0; k < n; ++k) {
        size_t o = (size_t)order[k];
        out[3 * k] = a[o]; out[3 * k + 1] = b[o]; out[3 * k + 2] = c[o];
      }
    } else if (field == ESPP_FIELD_IMAGE) {
      std::vector<int4> im = download(s->img, n, s->stream);
      for (size_t k = 0; k < n; ++k) {
        const int4& q = im[(size_t)order[k]];
        out[3 * k] = q.x; out[3 * k + 1] = q.y; out[3 * k + 2] = q.z;
      }
    } else {
      fail(ESPP_ERR_ARG, "unknown particle field");
    }
  });
}

// Slots of m particles given by id (the cell order must be current: slot_of_id is
// written by the binning).  Fails with the reference's message for an id that
// does not exist.
static void slots_of_ids(espp_sim* s, int32_t field, int64_t m, const int32_t* ids, double* out_values,
                         DBuf<int>& dslots) {
  s->ensure_sorted();
  DBuf<int> dids;
  DBuf<double> dout;
  dids.resize((size_t)m);
  dslots.resize((size_t)m);
  dout.resize((size_t)m * 3);
  hip_check(hipMemcpyAsync(dids.p, ids, (size_t)m * sizeof(int), hipMemcpyHostToDevice, s->stream), "id upload");
  launch_gather_ids(s->stream, (int)m, dids.p, s->slot_of_id.p, s->nid, field < 0 ? 0 : field, s->pos.p, s->vel.p,
                    s->fx.p, s->fy.p, s->fz.p, s->img.p, dout.p, dslots.p);
  std::vector<int> slots = download(dslots, (size_t)m, s->stream);
  for (int64_t k = 0; k < m; ++k)
    if (slots[(size_t)k] < 0) fail(ESPP_ERR_ARG, "particle " + std::to_string(ids[k]) + " does not exist");
  if (out_values) {
    std::vector<double> v = download(dout, (size_t)m * 3, s->stream);
    std::memcpy(out_values, v.data(), v.size() * sizeof(double));
  }
}

int espp_get_particles_by_id(espp_sim* s, int32_t field, int64_t m, const int32_t* ids, double* out) {
  return guarded([&] {
    need(s && (m == 0 || (ids && out)), "NULL argument");
    need(m >= 0, "negative count");
    need(field >= 0 && field <= ESPP_FIELD_MASS, "unknown particle field");
    if (m == 0) return;
    DBuf<int> slots;
    std::vector<double> v((size_t)m * 3);
    slots_of_ids(s, field, m, ids, v.data(), slots);
    const int width = (field == ESPP_FIELD_TYPE || field == ESPP_FIELD_MASS) ? 1 : 3;
    for (int64_t k = 0; k < m; ++k)
      for (int c = 0; c < width; ++c) out[(size_t)k * width + c] = v[(size_t)k * 3 + c];
  });
}

int espp_set_particles(espp_sim* s, int32_t field, int64_t m, const int32_t* ids, const double* values) {
  return guarded([&] {
    need(s && (m == 0 || (ids && values)), "NULL argument");
    need(m >= 0, "negative count");
    need(field == ESPP_FIELD_POS || field == ESPP_FIELD_VEL || field == ESPP_FIELD_TYPE || field == ESPP_FIELD_MASS,
         "field cannot be set");
    if (m == 0) return;
    const int width = (field == ESPP_FIELD_TYPE || field == ESPP_FIELD_MASS) ? 1 : 3;
    std::vector<double> v(values, values + (size_t)m * width);
    for (int64_t k = 0; k < m; ++k) {
      if (field == ESPP_FIELD_TYPE) {
        need(v[(size_t)k] >= 0.0, "particle types must be non-negative");
        v[(size_t)k] = std::floor(v[(size_t)k]);
        s->maxtype = std::max(s->maxtype, (int)v[(size_t)k]);
      } else if (field == ESPP_FIELD_MASS) {
        need(v[(size_t)k] > 0.0, "particle mass must be positive");
      } else if (field == ESPP_FIELD_POS) {
        for (int c = 0; c < 3; ++c) {
          // a non-finite coordinate is stored as given; the next binning reports the
          // lost particle as the reference's decompose does
          const double x = v[(size_t)k * 3 + c];
          if (std::isfinite(x)) s->coord_bound = std::max(s->coord_bound, std::fabs(x));
        }
      }
    }
    DBuf<int> slots;
    slots_of_ids(s, -1, m, ids, nullptr, slots);
    DBuf<double> dv;
    dv.resize(v.size());
    upload_bytes(dv.p, v.data(), v.size() * sizeof(double), s->stream, "upload");
    launch_scatter_slots(s->stream, (int)m, slots.p, field, dv.p, s->pos.p, s->vel.p);
    s->sync();
    // a new position or type: re-binning refreshes the cell order, the fixed-point
    // mirror and the lists
    if (field == ESPP_FIELD_POS || field == ESPP_FIELD_TYPE) s->sorted = false;
  });
}

int espp_vl_create(espp_sim* s, double cutoff, int32_t* vl) {
  return guarded([&] {
    need(s && vl, "NULL argument");
    need(cutoff >= 0.0, "cutoff must be non-negative");
    auto* v = new VerletListState();
    v->cut = cutoff;
    s->vls.push_back(v);
    try {
      s->ensure_sorted();
      s->build_vl(*v);
    } catch (...) {
      s->vls.pop_back();
      delete v;
      throw;
    }
    *vl = (int32_t)s->vls.size() - 1;
  });
}

static VerletListState& get_vl(espp_sim* s, int32_t vl) {
  need(s != nullptr, "NULL handle");
  need(vl >= 0 && vl < (int32_t)s->vls.size(), "no such verlet list");
  return *s->vls[(size_t)vl];
}

int espp_vl_exclude(espp_sim* s, int32_t vl, int64_t m, const int32_t* pairs) {
  return guarded([&] {
    VerletListState& v = get_vl(s, vl);
    need(m >= 0 && (m == 0 || pairs), "invalid exclusion list");
    for (int64_t k = 0; k < m; ++k) v.excl.emplace_back(pairs[2 * k], pairs[2 * k + 1]);
    v.has_excl = !v.excl.empty();
    v.ex_nid = -1;
    if (v.has_excl) s->upload_exclusions(v);
    s->ensure_sorted();
    s->build_vl(v);
  });
}

int espp_vl_rebuild(espp_sim* s, int32_t vl) {
  return guarded([&] {
    VerletListState& v = get_vl(s, vl);
    s->ensure_sorted();
    s->build_vl(v);
  });
}

int espp_vl_size(espp_sim* s, int32_t vl, int64_t* npairs) {
  return guarded([&] {
    VerletListState& v = get_vl(s, vl);
    need(npairs != nullptr, "NULL argument");
    s->ensure_lists();
    s->sync();
    std::vector<int> c = download(v.cnt, (size_t)s->n, s->stream);
    int64_t tot = 0;
    for (int x : c) tot += x;
    *npairs = tot / 2;
  });
}

int espp_vl_builds(espp_sim* s, int32_t vl, int64_t* builds) {
  return guarded([&] {
    VerletListState& v = get_vl(s, vl);
    need(builds != nullptr, "NULL argument");
    *builds = v.builds;
  });
}

int espp_vl_pairs(espp_sim* s, int32_t vl, int32_t* out, int64_t cap, int64_t* total) {
  return guarded([&] {
    VerletListState& v = get_vl(s, vl);
    need(total != nullptr && (cap == 0 || out != nullptr), "NULL argument");
    s->ensure_lists();
    s->sync();
    size_t n = (size_t)s->n;
    std::vector<int> c = download(v.cnt, n, s->stream), ids = download(s->id, n, s->stream);
    const size_t rowlen = v.tiled ? (size_t)v.cap / 2 : (size_t)v.cap;
    std::vector<int> nb = download(v.nbr, (size_t)((n + 63) / 64) * rowlen * 64, s->stream);
    std::vector<TileRun> runs;
    std::vector<int> nruns;
    if (v.tiled) {
      size_t ntiles = (n + (size_t)s->T - 1) / (size_t)s->T;
      runs = download(s->runs, ntiles * kMaxRuns, s->stream);
      nruns = download(s->nruns, ntiles, s->stream);
    }
    auto entry_slot = [&](size_t i, int q) -> int {
      if (!v.tiled) return nb[(i / 64) * rowlen * 64 + (size_t)q * 64 + (i % 64)];
      // 16-byte groups of eight entries, interleaved over the 64 particles of a wavefront
      const size_t grp = ((i / 64) * (size_t)(v.cap / 8) + (size_t)(q / 8)) * 64 + (i % 64);
      uint32_t w = (uint32_t)nb[grp * 4 + (size_t)((q / 2) % 4)];
      int e = (q & 1) ? (int)(w >> 16) : (int)(w & 0xFFFFu);
      size_t t = i / (size_t)s->T;
      const TileRun* sr = runs.data() + t * kMaxRuns;
      int r = 0;
      while (r + 1 < nruns[t] && sr[r + 1].lds_off <= e) ++r;
      return sr[r].slot_start + (e - sr[r].lds_off);
    };
    int64_t k = 0;
    for (size_t i = 0; i < n; ++i) {         // the list of slot i
      for (int q = 0; q < c[i]; ++q) {
        int j = entry_slot(i, q);
        if (ids[i] < ids[(size_t)j]) {
          if (k < cap) { out[2 * k] = ids[i]; out[2 * k + 1] = ids[(size_t)j]; }
          ++k;
        }
      }
    }
    *total = k;
  });
}

int espp_vl_cutoff(espp_sim* s, int32_t vl, double* cut) {
  return guarded([&] {
    VerletListState& v = get_vl(s, vl);
    need(cut != nullptr, "NULL argument");
    *cut = v.cut + s->skin;
  });
}

int espp_fpl_create(espp_sim* s, int32_t* fpl) {
  return guarded([&] {
    need(s && fpl, "NULL argument");
    s->fpls.push_back(new FixedPairListState());
    *fpl = (int32_t)s->fpls.size() - 1;
  });
}

int espp_fpl_add(espp_sim* s, int32_t fpl, int64_t m, const int32_t* pairs) {
  return guarded([&] {
    need(s != nullptr, "NULL handle");
    need(fpl >= 0 && fpl < (int32_t)s->fpls.size(), "no such fixed pair list");
    need(m >= 0 && (m == 0 || pairs), "invalid bond list");
    FixedPairListState& f = *s->fpls[(size_t)fpl];
    for (int64_t k = 0; k < m; ++k) {
      need(pairs[2 * k] >= 0 && pairs[2 * k + 1] >= 0, "bond ids must be non-negative");
      f.a.push_back(pairs[2 * k]);
      f.b.push_back(pairs[2 * k + 1]);
    }
    f.dirty = true;
  });
}

int espp_fpl_size(espp_sim* s, int32_t fpl, int64_t* m) {
  return guarded([&] {
    need(s && m, "NULL argument");
    need(fpl >= 0 && fpl < (int32_t)s->fpls.size(), "no such fixed pair list");
    *m = (int64_t)s->fpls[(size_t)fpl]->a.size();
  });
}

int espp_lj_create(espp_sim* s, int32_t vl, int32_t* ia) {
  return guarded([&] {
    get_vl(s, vl);
    need(ia != nullptr, "NULL argument");
    auto* a = new InteractionState();
    a->kind = 0;
    a->list = vl;
    s->ias.push_back(a);
    *ia = (int32_t)s->ias.size() - 1;
  });
}

static InteractionState& get_ia(espp_sim* s, int32_t ia) {
  need(s != nullptr, "NULL handle");
  need(ia >= 0 && ia < (int32_t)s->ias.size(), "no such interaction");
  return *s->ias[(size_t)ia];
}

int espp_lj_set(espp_sim* s, int32_t ia, int32_t t1, int32_t t2, double eps, double sigma, double cutoff,
                double shift) {
  return guarded([&] {
    InteractionState& a = get_ia(s, ia);
    need(a.kind == 0, "not a Lennard-Jones interaction");
    need(t1 >= 0 && t2 >= 0 && t1 < kMaxTypes && t2 < kMaxTypes, "particle type out of range");
    LJParam p{};
    // LennardJones::preset (src/interaction/LennardJones.hpp:79-87)
    double sig2 = sigma * sigma, sig6 = sig2 * sig2 * sig2;
    p.dff1 = 48.0 * eps * sig6 * sig6;
    p.dff2 = 24.0 * eps * sig6;
    p.drc2 = cutoff * cutoff;
    p.sig2 = sig2;
    p.eps4 = 4.0 * eps;
    if (shift != shift) {
      // setAutoShift (src/interaction/Potential.hpp:214-222)
      if (std::isinf(cutoff)) p.dshift = 0.0;
      else {
        double frac2 = sigma * sigma / p.drc2, frac6 = frac2 * frac2 * frac2;
        p.dshift = 4.0 * eps * (frac6 * frac6 - frac6);
      }
    } else {
      p.dshift = shift;
    }
    p.ff1 = (float)p.dff1;
    p.ff2 = (float)p.dff2;
    p.rc2 = std::isinf(p.drc2) ? INFINITY : (float)p.drc2;
    p.shift = (float)p.dshift;
    p.set = 1;
    a.tab[(size_t)t1 * kMaxTypes + t2] = p;
    a.tab[(size_t)t2 * kMaxTypes + t1] = p;
    a.maxset = std::max(a.maxset, std::max(t1, t2));
    a.dirty = true;
  });
}

int espp_fene_create(espp_sim* s, int32_t fpl, double K, double r0, double rMax, double cutoff, double shift,
                     int32_t* ia) {
  return guarded([&] {
    need(s && ia, "NULL argument");
    need(fpl >= 0 && fpl < (int32_t)s->fpls.size(), "no such fixed pair list");
    auto* a = new InteractionState();
    a->kind = 1;
    a->list = fpl;
    a->fp.K = K;
    a->fp.r0 = r0;
    a->fp.rMax = rMax;
    a->fp.rMaxSqr = rMax * rMax;
    a->fp.cutoffSqr = cutoff * cutoff;
    a->fp.shift = shift;
    s->ias.push_back(a);
    *ia = (int32_t)s->ias.size() - 1;
  });
}

int espp_tab_create(espp_sim* s, int32_t vl, int32_t* ia) {
  return guarded([&] {
    get_vl(s, vl);
    need(ia != nullptr, "NULL argument");
    auto* a = new InteractionState();
    a->kind = 2;
    a->list = vl;
    s->ias.push_back(a);
    *ia = (int32_t)s->ias.size() - 1;
  });
}

static std::shared_ptr<HostTable> make_table(int32_t itype, int64_t n, const double* r, const double* e,
                                             const double* f, double cutoff) {
  need(itype >= 1 && itype <= 3, "interpolation type must be 1 (linear), 2 (Akima) or 3 (cubic)");
  need(r && e && f, "NULL argument");
  need(n >= 2 && (itype != 2 || n >= 4), "illegal file for tabulated potential");
  need(n < (int64_t)1 << 24, "table too long");
  need(r[n - 1] > r[0], "illegal distance for entries");
  auto t = std::make_shared<HostTable>();
  t->itype = itype;
  t->r.assign(r, r + n);
  t->e.assign(e, e + n);
  t->f.assign(f, f + n);
  t->cutoff = cutoff;
  return t;
}

int espp_tab_set(espp_sim* s, int32_t ia, int32_t t1, int32_t t2, int32_t itype, int64_t n, const double* r,
                 const double* e, const double* f, double cutoff) {
  return guarded([&] {
    InteractionState& a = get_ia(s, ia);
    need(a.kind == 2, "not a VerletList Tabulated interaction");
    need(t1 >= 0 && t2 >= 0 && t1 < kMaxTypes && t2 < kMaxTypes, "particle type out of range");
    auto t = make_table(itype, n, r, e, f, cutoff);
    a.tables[(size_t)t1 * kMaxTypes + t2] = t;
    a.tables[(size_t)t2 * kMaxTypes + t1] = t;
    a.maxset = std::max(a.maxset, std::max(t1, t2));
    a.dirty = true;
  });
}

int espp_tabbond_create(espp_sim* s, int32_t fpl, int32_t itype, int64_t n, const double* r, const double* e,
                        const double* f, double cutoff, int32_t* ia) {
  return guarded([&] {
    need(s && ia, "NULL argument");
    need(fpl >= 0 && fpl < (int32_t)s->fpls.size(), "no such fixed pair list");
    auto t = make_table(itype, n, r, e, f, cutoff);
    auto* a = new InteractionState();
    a->kind = 3;
    a->list = fpl;
    a->tables[0] = t;
    s->ias.push_back(a);
    *ia = (int32_t)s->ias.size() - 1;
  });
}

int espp_get_rng_state(espp_sim* s, uint64_t* seed, uint64_t* calls) {
  return guarded([&] {
    need(s && seed && calls, "NULL argument");
    *seed = s->seed;
    *calls = s->therm_calls;
  });
}

int espp_set_rng_state(espp_sim* s, uint64_t seed, uint64_t calls) {
  return guarded([&] {
    need(s != nullptr, "NULL handle");
    s->seed = seed;
    s->therm_calls = calls;
  });
}

int espp_get_run_state(espp_sim* s, double* maxdist, int32_t* resort) {
  return guarded([&] {
    need(s && maxdist && resort, "NULL argument");
    *maxdist = s->maxDist;
    *resort = s->resortFlag ? 1 : 0;
  });
}

int espp_set_run_state(espp_sim* s, double maxdist, int32_t resort) {
  return guarded([&] {
    need(s != nullptr, "NULL handle");
    need(maxdist >= 0.0, "negative displacement");
    s->maxDist = maxdist;
    s->resortFlag = resort != 0;
    for (auto* a : s->ias) a->in_build = -1;   // the pruned lists' displacement bound is lost
  });
}

int espp_rdf(espp_sim* s, int32_t nbins, double* out) {
  return guarded([&] {
    need(s && out, "NULL argument");
    need(nbins > 0 && nbins <= (1 << 24), "number of bins out of range");
    s->ensure_sorted();
    DBuf<unsigned long long> counts;
    counts.resize((size_t)nbins);
    hip_check(hipMemsetAsync(counts.p, 0, (size_t)nbins * sizeof(unsigned long long), s->stream), "rdf clear");
    // dr from the y edge: RadialDistrF assumes a cubic box (src/analysis/RadialDistrF.cpp:57-58)
    const double dr = s->g.halfL[1] / (double)nbins;
    launch_rdf(s->stream, (int)s->n, s->pos.p, s->g, dr, nbins, counts.p);
    s->sync();
    std::vector<unsigned long long> c = download(counts, (size_t)nbins, s->stream);
    // normalisation of RadialDistrF.cpp:128-137; the factor is formed in long double as there (M_PIl)
    const double rho = (double)s->n / (s->g.L[0] * s->g.L[1] * s->g.L[2]);
    const double factor = (double)(2.0L * 3.141592653589793238462643383279502884L * (long double)dr *
                                   (long double)rho * 1.0L * (long double)s->n);
    for (int i = 0; i < nbins; ++i) {
      const double radius = (i + 0.5) * dr;
      out[i] = (double)c[(size_t)i] / (factor * (radius * radius + dr * dr / 12.0));
    }
  });
}

int espp_ia_energy(espp_sim* s, int32_t ia, double* e) {
  return guarded([&] {
    InteractionState& a = get_ia(s, ia);
    need(e != nullptr, "NULL argument");
    double o[8];
    s->interaction_sums(a, o);
    *e = o[0];
  });
}

int espp_ia_virial(espp_sim* s, int32_t ia, double* w, double* t6) {
  return guarded([&] {
    InteractionState& a = get_ia(s, ia);
    need(w != nullptr, "NULL argument");
    double o[8];
    s->interaction_sums(a, o);
    *w = o[1];
    if (t6) for (int q = 0; q < 6; ++q) t6[q] = o[2 + q];
  });
}

// espp_ia_energy / espp_ia_virial that allocates nothing: the per-workgroup
// partials and their reduction live in the device's static scratch.  Needs lists,
// tile plan and parameter tables current (any run or force computation leaves
// them so); otherwise ESPP_ERR_STATE, since bringing them up to date allocates.
int espp_ia_energy_noalloc(espp_sim* s, int32_t ia, double* energy, double* virial, double* tensor6) {
  return guarded([&] {
    InteractionState& a = get_ia(s, ia);
    need(energy != nullptr, "NULL argument");
    if (a.kind == 4) fail(ESPP_ERR_STATE, "the learned pair force is not conservative: it has no energy or virial");
    if (!s->scratch_dev) fail(ESPP_ERR_STATE, "no device scratch for this handle");
    const char* stale = "state needs a rebuild or upload first (call espp_compute_forces or espp_run once): "
                        "that allocates, this entry point does not";
    if (!s->sorted || (s->plan_pending) || (!s->plan_valid && s->n > 0)) fail(ESPP_ERR_STATE, stale);
    bool tiled = false;
    if (a.kind == 0 || a.kind == 2) {
      VerletListState& v = *s->vls[(size_t)a.list];
      if (v.stale) fail(ESPP_ERR_STATE, stale);
      tiled = v.tiled;
      if (a.dirty || a.dtab_nt != s->lj_ntypes(a)) fail(ESPP_ERR_STATE, stale);
    } else {
      FixedPairListState& f = *s->fpls[(size_t)a.list];
      if (f.dirty || f.built_nid != s->nid) fail(ESPP_ERR_STATE, stale);
      if (a.kind == 3 && (a.dirty || a.dtab_nt != 1)) fail(ESPP_ERR_STATE, stale);
    }
    const int n = (int)s->n;
    double out[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (n > 0) {
      const int nblk = std::max(tiled ? lj_tiled_blocks(n, s->T) : lj_blocks(n), 1);
      if ((size_t)nblk * 8 + 8 > kScratchDoubles)
        fail(ESPP_ERR_ARG, "system too large for the fixed scratch: " + std::to_string(nblk) +
                               " workgroups need " + std::to_string((size_t)nblk * 8 + 8) + " doubles, the scratch holds " +
                               std::to_string(kScratchDoubles));
      double* part = s->scratch_dev;
      double* tail = s->scratch_dev + kScratchDoubles - 8;
      s->run_interaction(a, true, false, part);
      launch_reduce_partials(s->stream, nblk, 8, part, tail);
      hip_check(hipMemcpyAsync(s->hred, tail, 8 * sizeof(double), hipMemcpyDeviceToHost, s->stream),
                "scratch readback");
      s->sync();
      std::memcpy(out, s->hred, sizeof(out));
    }
    *energy = out[0];
    if (virial) *virial = out[1];
    if (tensor6) for (int q = 0; q < 6; ++q) tensor6[q] = out[2 + q];
  });
}

int espp_add_interaction(espp_sim* s, int32_t ia) {
  return guarded([&] {
    get_ia(s, ia);
    for (int k : s->active) need(k != ia, "interaction already added");
    s->active.push_back(ia);
  });
}

int espp_remove_interaction(espp_sim* s, int32_t position) {
  return guarded([&] {
    need(s != nullptr, "NULL handle");
    if (position < 0 || position >= (int32_t)s->active.size())
      fail(ESPP_ERR_ARG, "Probably you are trying to remove the interaction " + std::to_string(position) +
                             " which does not exist. Check your script!");
    s->active.erase(s->active.begin() + position);
  });
}

int espp_fene_set(espp_sim* s, int32_t ia, double K, double r0, double rMax, double cutoff, double shift) {
  return guarded([&] {
    InteractionState& a = get_ia(s, ia);
    need(a.kind == 1, "not a FixedPairList FENE interaction");
    a.fp.K = K;
    a.fp.r0 = r0;
    a.fp.rMax = rMax;
    a.fp.rMaxSqr = rMax * rMax;
    a.fp.cutoffSqr = cutoff * cutoff;
    a.fp.shift = shift;
  });
}

int espp_ia_set_fpl(espp_sim* s, int32_t ia, int32_t fpl) {
  return guarded([&] {
    InteractionState& a = get_ia(s, ia);
    need(a.kind == 1 || a.kind == 3, "not a FixedPairList interaction");
    need(fpl >= 0 && fpl < (int32_t)s->fpls.size(), "no such fixed pair list");
    a.list = fpl;
  });
}

int espp_compute_forces(espp_sim* s) {
  return guarded([&] {
    need(s != nullptr, "NULL handle");
    s->calc_forces();
    s->sync();
  });
}

int espp_langevin(espp_sim* s, int32_t on, double gamma, double temperature, int64_t nexcl, const int32_t* excl) {
  return guarded([&] {
    need(s != nullptr, "NULL handle");
    need(nexcl >= 0 && (nexcl == 0 || excl), "invalid exclusion list");
    s->lang = on != 0;
    s->gamma = gamma;
    s->temperature = temperature;
    for (int64_t k = 0; k < nexcl; ++k) s->lang_excl.push_back(excl[k]);
    s->excl_dirty = true;
  });
}

int espp_run(espp_sim* s, int32_t nsteps, double dt) {
  return guarded([&] {
    need(s != nullptr, "NULL handle");
    need(nsteps >= 0, "nsteps must be non-negative");
    // MDIntegrator::setTimeStep (src/integrator/MDIntegrator.cpp:51-63)
    if (dt == 0.0) fail(ESPP_ERR_ARG, "Timestep 'dt' must be non-zero!");
    s->run(nsteps, dt);
  });
}

int espp_get_step(espp_sim* s, int64_t* step) {
  return guarded([&] {
    need(s && step, "NULL argument");
    *step = s->step;
  });
}

int espp_set_step(espp_sim* s, int64_t step) {
  return guarded([&] {
    need(s != nullptr, "NULL handle");
    s->step = step;
  });
}

int espp_num_resorts(espp_sim* s, int64_t* nr) {
  return guarded([&] {
    need(s && nr, "NULL argument");
    *nr = s->nResorts;
  });
}

int espp_timers(espp_sim* s, double out10[10]) {
  return guarded([&] {
    need(s && out10, "NULL argument");
    std::memcpy(out10, s->timers, sizeof(s->timers));
  });
}

static void kinetic_sums(espp_sim* s, unsigned long long mask, double out8[8]) {
  for (int q = 0; q < 8; ++q) out8[q] = 0.0;
  if (s->n == 0) return;
  int nblk = kinetic_blocks((int)s->n);
  s->partial.resize((size_t)nblk * 8);
  launch_kinetic(s->stream, (int)s->n, s->vel.p, s->pos.p, mask, s->partial.p);
  s->reduce(nblk, 8, out8);
}

int espp_kinetic(espp_sim* s, double* mv2) {
  return guarded([&] {
    need(s && mv2, "NULL argument");
    double o[8];
    kinetic_sums(s, ~0ull, o);
    *mv2 = o[0];
  });
}

int espp_kinetic_tensor(espp_sim* s, uint64_t type_mask, double* mv2, int64_t* count, double* tensor6) {
  return guarded([&] {
    need(s != nullptr, "NULL handle");
    double o[8];
    kinetic_sums(s, (unsigned long long)type_mask, o);
    if (mv2) *mv2 = o[0];
    if (count) *count = (int64_t)std::llround(o[1]);
    if (tensor6) for (int q = 0; q < 6; ++q) tensor6[q] = o[2 + q];
  });
}

// ------------------------------------------------------------------ instrumentation

int espp_device_synchronize(void) {
  return guarded([&] { hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize"); });
}

int espp_prof_enable(espp_sim* s, int32_t on) {
  return guarded([&] {
    need(s != nullptr, "NULL handle");
    s->prof_collect();
    s->prof = on != 0;
  });
}

int espp_prof_reset(espp_sim* s) {
  return guarded([&] {
    need(s != nullptr, "NULL handle");
    s->prof_collect();
    for (int k = 0; k < 8; ++k) { s->prof_count[k] = 0; s->prof_ms[k] = 0.0; }
  });
}

int espp_prof_stats(espp_sim* s, int32_t kernel, int64_t* launches, double* total_ms) {
  return guarded([&] {
    need(s && launches && total_ms, "NULL argument");
    need(kernel >= 0 && kernel < 8, "kernel id out of range");
    s->prof_collect();
    *launches = s->prof_count[kernel];
    *total_ms = s->prof_ms[kernel];
  });
}

int espp_set_step_window(espp_sim* s, int32_t max_steps) {
  return guarded([&] {
    need(s != nullptr, "NULL handle");
    need(max_steps >= 0 && max_steps <= 4096, "window length must be between 0 and 4096");
    s->spec_max = max_steps;
  });
}

int espp_step_window_stats(espp_sim* s, int64_t* windows, int64_t* launched, int64_t* ran) {
  return guarded([&] {
    need(s && windows && launched && ran, "NULL argument");
    *windows = s->spec_windows;
    *launched = s->spec_launched;
    *ran = s->spec_ran;
  });
}

int espp_set_prune_shell(espp_sim* s, double width) {
  return guarded([&] {
    need(s != nullptr, "NULL handle");
    need(std::isfinite(width), "prune shell must be finite");
    s->prune_shell = width;
    for (auto* a : s->ias) a->in_build = -1;
  });
}

int espp_prune_stats(espp_sim* s, int32_t ia, int64_t* prunes, int64_t* pruned_walks, int64_t* entries) {
  return guarded([&] {
    need(s && prunes && pruned_walks && entries, "NULL argument");
    need(ia >= 0 && ia < (int)s->ias.size(), "bad interaction handle");
    *prunes = s->prunes;
    *pruned_walks = s->pruned_walks;
    *entries = 0;
    InteractionState& a = *s->ias[(size_t)ia];
    if (a.kind == 0 && a.in_build >= 0 && a.in_cnt.n >= (size_t)s->n && s->n > 0) {
      std::vector<int> c = download(a.in_cnt, (size_t)s->n, s->stream);
      for (int x : c) *entries += x;
    }
  });
}

int espp_tile_stats(espp_sim* s, int32_t* tile, int32_t* max_region) {
  return guarded([&] {
    need(s != nullptr, "NULL handle");
    s->ensure_lists();
    s->ensure_plan();
    if (tile) *tile = s->T;
    if (max_region) *max_region = s->max_region;
  });
}

int espp_vl_stats(espp_sim* s, int32_t vl, int64_t* entries, int32_t* cap, int32_t* maxcount) {
  return guarded([&] {
    VerletListState& v = get_vl(s, vl);
    s->ensure_lists();
    s->sync();
    std::vector<int> c = download(v.cnt, (size_t)s->n, s->stream);
    int64_t tot = 0;
    for (int x : c) tot += x;
    if (entries) *entries = tot;
    if (cap) *cap = v.cap;
    if (maxcount) *maxcount = v.maxcount;
  });
}

int espp_vl_batch_stats(espp_sim* s, int32_t vl, int32_t batch, int64_t* lane_batches, int64_t* wave_batches) {
  return guarded([&] {
    VerletListState& v = get_vl(s, vl);
    if (batch < 1) fail(ESPP_ERR_ARG, "batch must be positive");
    s->ensure_lists();
    s->sync();
    std::vector<int> c = download(v.cnt, (size_t)s->n, s->stream);
    if (v.tiled) {
      // the pair kernels take the lists in lane order: lane l works on slot lane_slot[l]
      std::vector<int> ls = download(v.lane_slot, (size_t)s->n, s->stream), cl(c.size());
      for (size_t l = 0; l < c.size(); ++l) cl[l] = c[(size_t)ls[l]];
      c.swap(cl);
    }
    int64_t lanes = 0, waves = 0;
    for (size_t w0 = 0; w0 < c.size(); w0 += 64) {
      int64_t mx = 0;
      for (size_t l = w0; l < std::min(c.size(), w0 + 64); ++l) {
        const int64_t b = (c[l] + batch - 1) / batch;
        lanes += b;
        mx = std::max(mx, b);
      }
      waves += mx;
    }
    if (lane_batches) *lane_batches = lanes;
    if (wave_batches) *wave_batches = waves;
  });
}

}  // extern "C"

// ------------------------------------------------------------------ learned pair force
extern "C" {

int espp_mlp_create(espp_sim* s, int32_t vl, const double spec[7], int32_t* ia) {
  return guarded([&] {
    VerletListState& v = get_vl(s, vl);
    need(ia != nullptr && spec != nullptr, "NULL argument");
    const double rc = spec[0], Ra = spec[1], Rb = spec[2];
    need(rc > 0.0 && std::isfinite(rc), "the model cutoff must be positive and finite");
    need(Ra > 0.0 && Rb > 0.0 && Ra <= v.cut && Rb <= v.cut, "density radii must be positive and within the list cutoff");
    need(rc <= v.cut, "the model cutoff must not exceed the Verlet list cutoff");
    need(spec[4] > 0.0 && spec[6] > 0.0, "density scales must be positive");
    auto* a = new InteractionState();
    a->kind = 4;
    a->list = vl;
    mlp::Spec& sp = a->spec;
    sp.rc = (float)rc;
    sp.mu0 = 0.0f;
    sp.dmu = (float)(rc / (mlp::kRbf - 1));
    sp.inv2s2 = (float)(1.0 / (2.0 * (rc / (mlp::kRbf - 1)) * (rc / (mlp::kRbf - 1))));
    sp.R[0] = (float)Ra;
    sp.R[1] = (float)Rb;
    sp.mean[0] = (float)spec[3];
    sp.inv_scale[0] = (float)(1.0 / spec[4]);
    sp.mean[1] = (float)spec[5];
    sp.inv_scale[1] = (float)(1.0 / spec[6]);
    a->theta.resize(mlp::kParams);
    a->grad.resize(mlp::kParams);
    a->m1.resize(mlp::kParams);
    a->m2.resize(mlp::kParams);
    a->wpack.resize(mlp::kPackElems);
    a->loss.resize(1);
    mlp::launch_zero(s->stream, a->grad.p, mlp::kParams);
    mlp::launch_zero(s->stream, a->m1.p, mlp::kParams);
    mlp::launch_zero(s->stream, a->m2.p, mlp::kParams);
    hip_check(hipGetLastError(), "mlp create");
    v.untiled_only = true;
    v.stale = true;
    s->ias.push_back(a);
    *ia = (int32_t)s->ias.size() - 1;
  });
}

static InteractionState& get_mlp(espp_sim* s, int32_t ia) {
  InteractionState& a = get_ia(s, ia);
  if (a.kind != 4) fail(ESPP_ERR_ARG, "not a learned pair-force interaction");
  return a;
}

int espp_mlp_set_params(espp_sim* s, int32_t ia, const float* theta, int64_t n) {
  return guarded([&] {
    InteractionState& a = get_mlp(s, ia);
    need(theta != nullptr, "NULL argument");
    need(n == mlp::kParams, "the model has 10625 parameters");
    upload_bytes(a.theta.p, theta, (size_t)n * sizeof(float), s->stream, "mlp parameters");
    mlp::launch_pack(s->stream, a.theta.p, a.wpack.p);
    hip_check(hipGetLastError(), "mlp pack");
    a.params_set = true;
  });
}

int espp_mlp_get_params(espp_sim* s, int32_t ia, float* theta, int64_t n) {
  return guarded([&] {
    InteractionState& a = get_mlp(s, ia);
    need(theta != nullptr && n == mlp::kParams, "bad argument");
    auto h = download(a.theta, (size_t)n, s->stream);
    std::memcpy(theta, h.data(), (size_t)n * sizeof(float));
  });
}

int espp_mlp_grad(espp_sim* s, int32_t ia, float* out, int64_t n) {
  return guarded([&] {
    InteractionState& a = get_mlp(s, ia);
    need(out != nullptr && n == mlp::kParams, "bad argument");
    auto h = download(a.grad, (size_t)n, s->stream);
    std::memcpy(out, h.data(), (size_t)n * sizeof(float));
  });
}

int espp_mlp_zero_grad(espp_sim* s, int32_t ia) {
  return guarded([&] {
    InteractionState& a = get_mlp(s, ia);
    mlp::launch_zero(s->stream, a.grad.p, mlp::kParams);
    hip_check(hipGetLastError(), "mlp zero");
  });
}

int espp_mlp_train_frame(espp_sim* s, int32_t ia, const double* fref, double weight, double* loss) {
  return guarded([&] {
    InteractionState& a = get_mlp(s, ia);
    need(fref != nullptr, "NULL argument");
    if (!a.params_set) fail(ESPP_ERR_STATE, "the learned pair force has no parameters (setParameters)");
    int lds = 0;
    hip_check(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, s->device), "device attribute");
    if ((size_t)lds < mlp::lds_bytes(true) + 2048)
      fail(ESPP_ERR_DEVICE, "training needs " + std::to_string(mlp::lds_bytes(true) + 2048) +
                                " bytes of LDS per workgroup; the device offers " + std::to_string(lds));
    s->ensure_lists();
    VerletListState& v = *s->vls[(size_t)a.list];
    const int64_t n = s->n;
    // reference forces by id (input rows follow ascending id order)
    std::vector<int32_t> ids = s->host_ids;
    std::sort(ids.begin(), ids.end());
    std::vector<double> byid((size_t)std::max(s->nid, 1) * 3, 0.0);
    for (size_t k = 0; k < ids.size(); ++k)
      for (int c = 0; c < 3; ++c) byid[(size_t)ids[k] * 3 + c] = fref[3 * k + c];
    a.fref.resize(byid.size());
    upload_bytes(a.fref.p, byid.data(), byid.size() * sizeof(double), s->stream, "reference forces");
    a.rho.resize((size_t)std::max<int64_t>(n, 1));
    a.fpred.resize((size_t)std::max<int64_t>(n, 1) * 3);
    mlp::launch_density(s->stream, (int)n, s->pos.p, v.nbr.p, v.cnt.p, v.cap, s->g, a.spec, a.rho.p);
    mlp::launch_forces(s->stream, (int)n, s->pos.p, v.nbr.p, v.cnt.p, v.cap, s->g, a.spec, a.rho.p, a.wpack.p,
                       a.theta.p, false, s->fx.p, s->fy.p, s->fz.p);
    mlp::launch_zero_d(s->stream, a.loss.p);
    const double scale = n > 0 ? weight / (3.0 * (double)n) : 0.0;
    mlp::launch_train(s->stream, (int)n, s->pos.p, s->id.p, v.nbr.p, v.cnt.p, v.cap, s->g, a.spec, a.rho.p, a.wpack.p,
                      a.theta.p, s->fx.p, s->fy.p, s->fz.p, a.fpred.p, a.fref.p, scale, a.grad.p, a.loss.p);
    hip_check(hipGetLastError(), "mlp train");
    auto h = download(a.loss, 1, s->stream);
    if (loss) *loss = h[0];
  });
}

int espp_mlp_adam(espp_sim* s, int32_t ia, double lr, double beta1, double beta2, double eps) {
  return guarded([&] {
    InteractionState& a = get_mlp(s, ia);
    need(lr > 0.0 && beta1 >= 0.0 && beta1 < 1.0 && beta2 >= 0.0 && beta2 < 1.0 && eps > 0.0,
         "Adam needs lr > 0, 0 <= beta < 1, eps > 0");
    if (!a.params_set) fail(ESPP_ERR_STATE, "the learned pair force has no parameters (setParameters)");
    a.adam_steps++;
    const double c1 = 1.0 / (1.0 - std::pow(beta1, (double)a.adam_steps));
    const double c2 = 1.0 / (1.0 - std::pow(beta2, (double)a.adam_steps));
    mlp::launch_adam(s->stream, a.theta.p, a.grad.p, a.m1.p, a.m2.p, (float)lr, (float)beta1, (float)beta2,
                     (float)eps, (float)c1, (float)c2);
    mlp::launch_pack(s->stream, a.theta.p, a.wpack.p);
    hip_check(hipGetLastError(), "mlp adam");
  });
}

}  // extern "C"
