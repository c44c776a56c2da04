// flow.hip -- continuous-batching server of O(N) LPA flows with an on-device RK4 stepper.
//
// Each flow evolves u(rho) = U'(rho) on a uniform rho grid from k = Lambda = 1 to k_end with fixed
// RK4 steps in t = ln(Lambda/k):
//   du/dt = 1/2 [ (N-1) u' I_T(u) + (3 u' + 2 rho u'') I_T(u + 2 rho u') ],
//   I_T(m2) = int d^3q/(2pi)^3 RBdot(k^2,q^2) S2(E),  E^2 = q^2 + RB(k^2,q^2) + m2,
//   S2(E) = T sum_n (w_n^2 + E^2)^-2 = coth(E/2T)/(4E^3) + csch^2(E/2T)/(8 T E^2),
// i.e. two threshold integrals (the d = 3 radial form of IntegratorGPU<3>, integrator_gpu.hh:34-60,
// with the closed-form bosonic Matsubara sum, cf. Examples/ONfiniteT/flows/V/V.kernel) per grid
// point and RK stage.  u', u'' are second-order finite differences (one-sided at the ends).
//
// Device layout: every resident flow owns 2 n doubles in one HBM pool (u after even and after odd
// steps) and a slot in SoA descriptor arrays.  The host cuts the resident flows into runs of
// consecutive grid points; one launch advances every run by a whole RK4 step, one workgroup per
// run, evaluating the stages outside the run as far as the next stages' stencils reach
// (flow_step_kernel), each wave two grid points at a time (one Gauss-Legendre node per lane, the
// four node sums reduced together across lanes).  steps_per_launch steps are captured into one hipGraph.  A flow's step
// index is the launch's base step plus the step within the graph minus the base step at its
// admission, so flows admitted at different times share the graph and a finished flow simply
// stops updating.  The graph's last node advances the device copy
// of the base step, which the host mirrors.  Between graph launches the host retires
// finished and non-finite flows and admits queued ones while their HBM footprint fits the budget.
//
// Node factors: with x_j the GL nodes, xr_j = x_j x_extent, the regulator terms are
// q^2 = k^2 xr, R = k^2 e^-xr, RBdot = 2 k^2 e^-xr (1 + xr) and the d = 3 element times weight is
// k^3 sqrt(xr) w_j x_extent / (4 pi^2), so per node only A_j = xr_j + e^-xr_j (E^2 = k^2 A_j + m^2)
// and D_j = sqrt(xr_j) w_j x_extent e^-xr_j (1 + xr_j) / (2 pi^2) (element * weight * RBdot = k^5 D_j)
// depend on j; both are fixed by the quadrature and tabulated once.  k(t) is tabulated per step and
// stage exactly as the oracle's stepper forms it.
#include "internal.hh"
#include "dfrg/device.hh"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <cmath>
#include <deque>
#include <map>
#include <vector>

using namespace dfrg;

namespace
{
  constexpr int kFlowGrid = 1024; // workgroups per step launch, grid-stride over runs
  constexpr double kPi = 3.14159265358979323846;

  // A run of consecutive grid points [a, b) of one resident flow, which one workgroup advances by
  // a whole RK4 step.  Runs are at least 2 points long (the one-sided stencils at the grid ends
  // then find their points in the previous stage's range) and at most kMaxRun.
  struct FlowRun {
    int slot, a, b, n, off, admit;
  };

  constexpr int kRunBlock = 512;                  // threads per workgroup: 8 waves share a run
  constexpr int kHalo = 3;                        // stages 1-3 are also evaluated this far outside the run
  constexpr int kMaxRun = 256;                    // longest run
  constexpr int kSpan = kMaxRun + 2 * (kHalo + 1); // the run and the stencil reach of its stage 1
  // grid points a wave evaluates together: 2 (rather than 4) balances a stage's points over the 8
  // waves at short runs; each node sum gets the same add tree either way (dev::wave_sum4)
  constexpr int kGroup = 2;

  struct FlowDev {
    int *status, *bad;    // per slot; status 0 idle, 1 running (host-owned)
    double *h, *T, *N;    // per slot: grid spacing, temperature, O(N)
    double *m2_0, *lam;   // per slot: initial condition u(rho) = m2_0 + lam/2 rho (step 0 reads it)
    FlowRun *runs;        // every resident flow cut into runs
    int *n_runs;          // device copy of the run count
    int *base;            // global step index of the first step of this launch
    double *pool;         // state: u at even steps | u at odd steps, per flow
    const double *A, *D;  // node factors (header comment)
    const double *kt;     // k at (step, stage point): 3 per step, t, t + dt/2, t + dt
    int nx, n_steps;
    double dt;
  };
  // S2(E) = coth(E/2T)/(4E^3) + csch^2(E/2T)/(8 T E^2) with e = exp(-E/T), coth = (1+e)/(1-e),
  // csch^2 = 4e/(1-e)^2, put over one denominator:
  //   S2 = (2 T (1+e)(1-e) + 4 e E) / (8 T E^3 (1-e)^2),   E = sqrt(E2).
  // FP64-accurate seeded forms instead of library calls: y = 1/E from an FP32 rsqrt seed and one
  // Newton step, E = E2 y; e from the 1/64-grid table and a degree-6 polynomial; the denominator as
  // the product of 1/(8T), y^3 and the seeded reciprocal of (1-e), squared, so no factor leaves the
  // FP32 range of the seeds.  1 - e loses relative accuracy only as E/T -> 0 (absolute error
  // ~1e-16 against 1 - e ~ E/T).  Relative error of S2 ~1e-14.
  __device__ __forceinline__ double bose_S2(double E2, double T, double inv_8T, double inv_T,
                                            const double *exp2t)
  {
    const double y = dev::rsqrt_seeded(E2);
    const double E = E2 * y;
    const double e = dev::exp_neg_tab(E * inv_T, exp2t), ome = 1.0 - e;
    const double num = fma(2.0 * T * (1.0 + e), ome, 4.0 * e * E);
    const double r = dev::rcp_seeded(ome);
    return num * inv_8T * (y * y * y) * (r * r);
  }

  // One RK4 step of every resident flow.  A workgroup takes a run [a, b) of one flow and keeps the
  // flow's values at the points [a - kHalo - 1, b + kHalo] in LDS.  Stage s (halo w = 4 - s)
  // evaluates the right-hand side at the points of [a - w, b + w) inside the grid, from the
  // previous stage's values at their neighbours, which the previous stage computed on a range one
  // point wider: points outside the run are evaluated again by the workgroup that owns them, with
  // the same inputs and arithmetic, so every point's result is what a stage-by-stage sweep gives
  // and does not depend on how flows are cut into runs.  Within a stage the waves share the points
  // two at a time (lanes over Gauss-Legendre nodes).  u is read from the buffer of this step's
  // parity and written to the other, so neighbouring runs read the old u.
  __global__ __launch_bounds__(kRunBlock) void flow_step_kernel(FlowDev F, int local)
  {
    __shared__ double s_exp2[64]; // 2^(j/64)
    __shared__ double s_u[kSpan], s_y[kSpan], s_acc[kSpan], s_f[kSpan];
    __shared__ double s_m2pi[kSpan], s_m2sig[kSpan], s_up[kSpan], s_upp[kSpan];
    for (int j = threadIdx.x; j < 64; j += kRunBlock) s_exp2[j] = exp2(j * (1.0 / 64.0));
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int n_wv = kRunBlock / 64;
    const int R = *F.n_runs;
    const int base = *F.base + local;
    // node factors, one node per lane when there are at most 64 (the rest loop over nodes)
    const bool lane_nodes = F.nx <= 64;
    double A_l = 0.0, D_l = 0.0;
    if (lane_nodes && lane < F.nx) {
      A_l = F.A[lane];
      D_l = F.D[lane];
    }
    __syncthreads();
    for (int r = blockIdx.x; r < R; r += gridDim.x) {
      const FlowRun run = F.runs[r]; // workgroup-uniform
      const int step = base - run.admit;
      // a flow outside its steps, or idle, is left alone (workgroup-uniform)
      if (step < 0 || step >= F.n_steps || F.status[run.slot] != 1) continue;
      const int n = run.n, a = run.a, b = run.b;
      const double h = F.h[run.slot], T = F.T[run.slot], Nn = F.N[run.slot], dt = F.dt;
      const double k_t = F.kt[3 * step], k_half = F.kt[3 * step + 1], k_dt = F.kt[3 * step + 2];
      const double inv_T = 1.0 / T, inv_8T = 0.125 * inv_T;
      const double *u_old = F.pool + run.off + (step & 1) * n;
      double *u_new = F.pool + run.off + ((step + 1) & 1) * n;
      const int x0 = a - kHalo - 1; // grid point of LDS entry 0
      // step 0 starts from the initial condition itself, so admission copies nothing to the device
      const bool initial = step == 0;
      const double m2_0 = F.m2_0[run.slot], half_lam = 0.5 * F.lam[run.slot];
      for (int i = tid; i < b - a + 2 * (kHalo + 1); i += kRunBlock) {
        const int x = x0 + i;
        const double v = x >= 0 && x < n ? (initial ? m2_0 + half_lam * (x * h) : u_old[x]) : 0.0;
        s_u[i] = v;
        s_y[i] = v;
      }
      __syncthreads();
      for (int stage = 1; stage <= 4; ++stage) {
        const int w = 4 - stage;
        const int p0 = max(a - w, 0) - x0, p1 = min(b + w, n) - x0; // evaluated entries [p0, p1)
        const double k = stage == 1 ? k_t : (stage == 4 ? k_dt : k_half);
        // stencil terms of the evaluated points from this stage's input
        for (int p = p0 + tid; p < p1; p += kRunBlock) {
          const int x = x0 + p;
          const double yv = s_y[p];
          double up, upp;
          if (x == 0) {
            up = (-3.0 * yv + 4.0 * s_y[p + 1] - s_y[p + 2]) / (2.0 * h);
            upp = (yv - 2.0 * s_y[p + 1] + s_y[p + 2]) / (h * h);
          } else if (x == n - 1) {
            up = (3.0 * yv - 4.0 * s_y[p - 1] + s_y[p - 2]) / (2.0 * h);
            upp = (yv - 2.0 * s_y[p - 1] + s_y[p - 2]) / (h * h);
          } else {
            up = (s_y[p + 1] - s_y[p - 1]) / (2.0 * h);
            upp = (s_y[p + 1] - 2.0 * yv + s_y[p - 1]) / (h * h);
          }
          const double rho = x * h;
          s_up[p] = up;
          s_upp[p] = upp;
          s_m2pi[p] = yv;
          s_m2sig[p] = yv + 2.0 * rho * up;
        }
        __syncthreads();
        // right-hand sides, two points per wave at a time: their four node sums (two threshold
        // integrals each) are reduced together by dev::wave_sum4 (kind of point pp: vector 2 pp + kind)
        const double k2 = k * k, k5 = k * k * (k * k) * k;
        for (int g0 = p0 + kGroup * wave; g0 < p1; g0 += kGroup * n_wv) {
          double v[2 * kGroup];
#pragma unroll
          for (int pp = 0; pp < kGroup; ++pp) {
            const int p = g0 + pp;
            double ipi = 0.0, isg = 0.0;
            if (p < p1) { // wave-uniform
              const double pm2pi = s_m2pi[p], pm2sig = s_m2sig[p];
              // two threshold integrals, lane j evaluates node j (d = 3 radial element)
              if (lane_nodes) {
                if (lane < F.nx) {
                  ipi = fma(D_l, bose_S2(fma(k2, A_l, pm2pi), T, inv_8T, inv_T, s_exp2), 0.0);
                  isg = fma(D_l, bose_S2(fma(k2, A_l, pm2sig), T, inv_8T, inv_T, s_exp2), 0.0);
                }
              } else {
                for (int j = lane; j < F.nx; j += 64) {
                  const double an = F.A[j], dd = F.D[j];
                  ipi = fma(dd, bose_S2(fma(k2, an, pm2pi), T, inv_8T, inv_T, s_exp2), ipi);
                  isg = fma(dd, bose_S2(fma(k2, an, pm2sig), T, inv_8T, inv_T, s_exp2), isg);
                }
              }
            }
            v[2 * pp] = ipi;
            v[2 * pp + 1] = isg;
          }
          const double sum = dev::wave_sum4(v);
#pragma unroll
          for (int pp = 0; pp < kGroup; ++pp) {
            const int p = g0 + pp;
            if (p >= p1) continue; // wave-uniform
            const double ipi = k5 * dev::lane_value(sum, dev::sum8_lane(2 * pp));
            const double isg = k5 * dev::lane_value(sum, dev::sum8_lane(2 * pp + 1));
            if (lane == 0) {
              const double up = s_up[p], upp = s_upp[p], rho = (x0 + p) * h;
              s_f[p] = 0.5 * ((Nn - 1.0) * up * ipi + (3.0 * up + 2.0 * rho * upp) * isg);
            }
          }
        }
        __syncthreads();
        // the RK4 stage update of the evaluated points
        for (int p = p0 + tid; p < p1; p += kRunBlock) {
          const double f = s_f[p], u = s_u[p];
          if (stage == 1) {
            s_acc[p] = f;
            s_y[p] = u + 0.5 * dt * f;
          } else if (stage == 2) {
            s_acc[p] = s_acc[p] + 2.0 * f;
            s_y[p] = u + 0.5 * dt * f;
          } else if (stage == 3) {
            s_acc[p] = s_acc[p] + 2.0 * f;
            s_y[p] = u + dt * f;
          } else {
            const double un = u + dt / 6.0 * (s_acc[p] + f);
            u_new[x0 + p] = un;
            if (!isfinite(un)) atomicOr(&F.bad[run.slot], 1);
          }
        }
        __syncthreads();
      }
    }
  }

  // last node of each graph launch: the next launch starts spl steps later
  __global__ void flow_base_kernel(int *base, int spl)
  {
    if (threadIdx.x == 0) *base += spl;
  }

  double now_s()
  {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  }
} // namespace

struct FlowRecord {
  dfrg_flow_desc desc;
  int state = 0; // 0 queued, 1 resident, 2 finished, 3 failed
  double t_submit = 0, t_done = 0;
  int slot = -1, off = -1, admit = 0;
  std::vector<double> u;
};

// The per-slot parameters and the run count live in one device block with a pinned host mirror:
// the host edits the mirror and uploads it with one asynchronous copy before a launch, and the
// launch's non-finite flags come back into the mirror behind the graph.
struct SlotLayout {
  size_t n_runs = 0, status = 0, bad = 0, h = 0, T = 0, N = 0, m2_0 = 0, lam = 0, bytes = 0;
};

struct dfrg_fs {
  size_t budget = 0;
  int max_resident = 0, nx = 64, n_steps = 2000, steps_per_launch = 10;
  double k_end = 1e-3, x_extent = 20.0;
  FlowDev dev{};
  size_t pool_doubles = 0;
  std::vector<std::pair<int, int>> free_ranges; // (offset, length) in doubles, sorted by offset
  std::vector<long long> slot_flow;             // flow id per slot, -1 idle
  std::map<long long, FlowRecord> flows;
  std::deque<long long> queue;
  long long next_id = 1;
  bool runs_dirty = true;
  size_t runs_cap = 0; // FlowRun entries allocated on the device
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int base = 0;           // global step index of the next launch's first step (mirrors dev.base)
  double *d_nodes = nullptr; // A | D | kt
  std::vector<double> latencies;
  double busy_s = 0;
  int finished = 0;
  // graph launches: HIP events around each on the server's stream, their summed device time, and
  // the grid-point RK4 steps they advanced
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  long long launches = 0, point_steps = 0;
  double device_s = 0;
  // slot block (device) and its pinned mirror; pinned staging of the run table and of the final u
  // of the flows that end in a launch (copied back behind the graph)
  SlotLayout sl;
  char *d_slots = nullptr, *h_slots = nullptr;
  bool slots_dirty = true;
  FlowRun *h_runs = nullptr;
  size_t h_runs_cap = 0;
  int runs_upload = 0;
  double *h_u = nullptr;
  size_t h_u_cap = 0;
  std::vector<std::pair<int, size_t>> u_staged; // (slot, offset in h_u) of this launch's finishers
  template <class T> T *mirror(size_t off) { return (T *)(h_slots + off); }
};

namespace
{
  int fs_alloc_range(dfrg_fs *fs, int len)
  {
    for (size_t r = 0; r < fs->free_ranges.size(); ++r) {
      auto &fr = fs->free_ranges[r];
      if (fr.second >= len) {
        const int off = fr.first;
        fr.first += len;
        fr.second -= len;
        if (fr.second == 0) fs->free_ranges.erase(fs->free_ranges.begin() + r);
        return off;
      }
    }
    return -1;
  }

  void fs_free_range(dfrg_fs *fs, int off, int len)
  {
    fs->free_ranges.push_back({off, len});
    std::sort(fs->free_ranges.begin(), fs->free_ranges.end());
    std::vector<std::pair<int, int>> merged;
    for (auto &r : fs->free_ranges) {
      if (!merged.empty() && merged.back().first + merged.back().second == r.first)
        merged.back().second += r.second;
      else
        merged.push_back(r);
    }
    fs->free_ranges.swap(merged);
  }

  int fs_capture(dfrg_fs *fs)
  {
    if (fs->graph_exec) return DFRG_OK;
    hipStream_t cs;
    DFRG_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    DFRG_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    for (int st = 0; st < fs->steps_per_launch; ++st)
      hipLaunchKernelGGL(flow_step_kernel, dim3(kFlowGrid), dim3(kRunBlock), 0, cs, fs->dev, st);
    hipLaunchKernelGGL(flow_base_kernel, dim3(1), dim3(64), 0, cs, fs->dev.base, fs->steps_per_launch);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(cs, &g);
    hipStreamDestroy(cs);
    if (e != hipSuccess) {
      set_error("flow graph capture failed: %s", hipGetErrorString(e));
      return DFRG_ERR_HIP;
    }
    fs->graph = g;
    DFRG_HIP(hipGraphInstantiate(&fs->graph_exec, g, nullptr, nullptr, 0));
    return DFRG_OK;
  }

  // retire finished / failed flows: copy u back, free their slot and HBM range
  int fs_retire(dfrg_fs *fs)
  {
    const int *bad = fs->mirror<int>(fs->sl.bad); // copied back behind the launch
    const double t = now_s();
    for (int s = 0; s < fs->max_resident; ++s) {
      if (fs->slot_flow[s] < 0) continue;
      FlowRecord &r = fs->flows[fs->slot_flow[s]];
      const bool done = fs->base - r.admit >= fs->n_steps;
      if (!done && !bad[s]) continue;
      const int n = r.desc.n_grid;
      r.u.resize(n);
      // u after the last step taken sits in the buffer of that step count's parity
      const int taken = std::min(fs->base - r.admit, fs->n_steps);
      const double *staged = nullptr;
      if (!bad[s])
        for (const auto &st : fs->u_staged)
          if (st.first == s) staged = fs->h_u + st.second;
      if (staged)
        std::copy(staged, staged + n, r.u.begin());
      else // a flow evicted as non-finite: its u after the last step taken
        DFRG_HIP(hipMemcpy(r.u.data(), fs->dev.pool + r.off + (taken & 1) * n, sizeof(double) * n,
                           hipMemcpyDeviceToHost));
      r.state = bad[s] ? 3 : 2;
      r.t_done = t;
      if (r.state == 2) {
        fs->latencies.push_back(r.t_done - r.t_submit);
        ++fs->finished;
      }
      fs_free_range(fs, r.off, 2 * n);
      fs->mirror<int>(fs->sl.status)[s] = 0;
      fs->slots_dirty = true;
      fs->slot_flow[s] = -1;
      r.slot = -1;
      fs->runs_dirty = true;
    }
    return DFRG_OK;
  }

  int fs_admit(dfrg_fs *fs)
  {
    while (!fs->queue.empty()) {
      const long long id = fs->queue.front();
      FlowRecord &r = fs->flows[id];
      const int n = r.desc.n_grid;
      int slot = -1;
      for (int s = 0; s < fs->max_resident; ++s)
        if (fs->slot_flow[s] < 0) {
          slot = s;
          break;
        }
      if (slot < 0) break;
      const int off = fs_alloc_range(fs, 2 * n);
      if (off < 0) break; // does not fit the HBM budget now: wait for retirements (FIFO)
      fs->queue.pop_front();
      // the slot's parameters go to the mirror; step 0 reads the initial condition from them, and
      // writes the odd buffer, so the flow's HBM range needs no initialisation
      const SlotLayout &L = fs->sl;
      fs->mirror<int>(L.status)[slot] = 1;
      fs->mirror<int>(L.bad)[slot] = 0;
      fs->mirror<double>(L.h)[slot] = r.desc.rho_max / (n - 1);
      fs->mirror<double>(L.T)[slot] = r.desc.T;
      fs->mirror<double>(L.N)[slot] = r.desc.N;
      fs->mirror<double>(L.m2_0)[slot] = r.desc.m2_0;
      fs->mirror<double>(L.lam)[slot] = r.desc.lambda;
      fs->slots_dirty = true;
      fs->slot_flow[slot] = id;
      r.slot = slot;
      r.off = off;
      r.admit = fs->base;
      r.state = 1;
      fs->runs_dirty = true;
    }
    if (fs->runs_dirty) {
      // cut every resident flow into runs of at most L points, with the smallest L (>= 8) whose
      // runs fit the workgroups resident at once (two per CU), so that a step takes one round of
      // workgroups (a flow of n >= 4 points then has runs of >= 2 points)
      std::vector<int> sizes;
      long long P = 0;
      for (int s = 0; s < fs->max_resident; ++s)
        if (fs->slot_flow[s] >= 0) {
          sizes.push_back(fs->flows[fs->slot_flow[s]].desc.n_grid);
          P += sizes.back();
        }
      const long long slots = 2LL * (context().cu_count > 0 ? context().cu_count : 256);
      int L = (int)std::min<long long>(kMaxRun, std::max<long long>(8, (P + slots - 1) / slots));
      auto count = [&](int len) {
        long long c = 0;
        for (int n : sizes) c += (n + len - 1) / len;
        return c;
      };
      while (L < kMaxRun && count(L) > slots) ++L;
      std::vector<FlowRun> runs;
      for (int s = 0; s < fs->max_resident; ++s) {
        if (fs->slot_flow[s] < 0) continue;
        const FlowRecord &r = fs->flows[fs->slot_flow[s]];
        const int n = r.desc.n_grid, cnt = (n + L - 1) / L, len = n / cnt, extra = n % cnt;
        for (int j = 0, a = 0; j < cnt; ++j) {
          const int b = a + len + (j < extra ? 1 : 0);
          runs.push_back(FlowRun{s, a, b, n, r.off, r.admit});
          a = b;
        }
      }
      const int R = (int)runs.size();
      if ((size_t)R > fs->runs_cap) {
        set_error("internal: %d flow runs exceed the %zu allocated", R, fs->runs_cap);
        return DFRG_ERR_RUNTIME;
      }
      // staged in pinned memory and uploaded before the next launch (no copy is in flight here:
      // every launch ends with a synchronisation)
      if ((size_t)R > fs->h_runs_cap) {
        const size_t cap = std::max<size_t>((size_t)R, 2 * fs->h_runs_cap);
        if (fs->h_runs) (void)hipHostFree(fs->h_runs);
        fs->h_runs = nullptr;
        fs->h_runs_cap = 0;
        DFRG_HIP(hipHostMalloc((void **)&fs->h_runs, sizeof(FlowRun) * cap, hipHostMallocDefault));
        fs->h_runs_cap = cap;
      }
      if (R) std::copy(runs.begin(), runs.end(), fs->h_runs);
      fs->runs_upload = R;
      *fs->mirror<int>(fs->sl.n_runs) = R;
      fs->slots_dirty = true;
      fs->runs_dirty = false;
    }
    return DFRG_OK;
  }

  int resident_count(const dfrg_fs *fs)
  {
    int c = 0;
    for (long long id : fs->slot_flow) c += id >= 0;
    return c;
  }
} // namespace

extern "C" {

size_t dfrg_fs_flow_bytes(int n_grid) { return (size_t)2 * n_grid * sizeof(double); }

int dfrg_fs_timing(dfrg_fs *fs, long long *launches, double *device_s, long long *point_steps)
{
  DFRG_CHECK_ARG(fs != nullptr, "NULL server");
  if (launches) *launches = fs->launches;
  if (device_s) *device_s = fs->device_s;
  if (point_steps) *point_steps = fs->point_steps;
  return DFRG_OK;
}

int dfrg_fs_create(dfrg_fs **out, dfrg_qp *qp, size_t hbm_budget, int max_resident, int nx, int n_steps,
                   double k_end, int steps_per_launch)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(out && qp, "NULL handle");
  DFRG_CHECK_ARG(max_resident >= 1 && max_resident <= 4096, "max_resident %d out of range [1,4096]", max_resident);
  DFRG_CHECK_ARG(nx >= 1 && nx <= 1024, "nx %d out of range", nx);
  DFRG_CHECK_ARG(n_steps >= 1 && k_end > 0 && k_end < 1, "invalid n_steps / k_end");
  DFRG_CHECK_ARG(steps_per_launch >= 1 && steps_per_launch <= 1000, "steps_per_launch out of range");
  const size_t max_points = (size_t)max_resident * 4096;
  DFRG_CHECK_ARG(hbm_budget >= dfrg_fs_flow_bytes(4) && hbm_budget / sizeof(double) < (size_t)1 << 31,
                 "hbm_budget out of range");
  dfrg_fs *fs = new dfrg_fs;
  fs->budget = hbm_budget;
  fs->max_resident = max_resident;
  fs->nx = nx;
  fs->n_steps = n_steps;
  fs->k_end = k_end;
  fs->steps_per_launch = steps_per_launch;
  fs->pool_doubles = hbm_budget / sizeof(double);
  fs->free_ranges.push_back({0, (int)fs->pool_doubles});
  fs->slot_flow.assign(max_resident, -1);
  FlowDev &d = fs->dev;
  d.nx = nx;
  d.n_steps = n_steps;
  d.dt = std::log(1.0 / k_end) / n_steps;
  const void *hx = nullptr, *hw = nullptr;
  int rc = dfrg_qp_gl(qp, nx, 0, nullptr, nullptr, &hx, &hw);
  if (rc != DFRG_OK) {
    delete fs;
    return rc;
  }
  // node factors and k(t) (header comment); k as the oracle's stepper forms it: t = s dt,
  // exp(-t), exp(-(t + dt/2)), exp(-(t + dt))
  std::vector<double> nodes(2 * nx + 3 * (size_t)n_steps);
  for (int j = 0; j < nx; ++j) {
    const double xr = ((const double *)hx)[j] * fs->x_extent;
    const double e = std::exp(-xr);
    nodes[j] = xr + e;
    nodes[nx + j] = std::sqrt(xr) * (((const double *)hw)[j] * fs->x_extent) * e * (1.0 + xr) / (2.0 * kPi * kPi);
  }
  for (int st = 0; st < n_steps; ++st) {
    const double t = st * d.dt;
    nodes[2 * nx + 3 * st] = std::exp(-t);
    nodes[2 * nx + 3 * st + 1] = std::exp(-(t + 0.5 * d.dt));
    nodes[2 * nx + 3 * st + 2] = std::exp(-(t + d.dt));
  }
  // runs: at most one per 4 points of each resident flow (n >= 4, L >= 8), plus one per flow
  fs->runs_cap = std::min(max_points, fs->pool_doubles / 4 + (size_t)max_resident + 1);
  const size_t runs_bytes = sizeof(FlowRun) * fs->runs_cap;
#define FS_ALLOC(ptr, bytes)                                                                                   \
  do {                                                                                                         \
    if (hipMalloc((void **)&(ptr), (bytes)) != hipSuccess || hipMemset((ptr), 0, (bytes)) != hipSuccess) {     \
      set_error("flow server: device allocation of %zu bytes failed", (size_t)(bytes));                        \
      dfrg_fs_destroy(fs);                                                                                     \
      return DFRG_ERR_HIP;                                                                                     \
    }                                                                                                          \
  } while (0)
  {
    SlotLayout &L = fs->sl;
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t M = (size_t)max_resident;
    size_t o = 0;
    L.n_runs = o, o += up(sizeof(int));
    L.status = o, o += up(sizeof(int) * M);
    L.bad = o, o += up(sizeof(int) * M);
    L.h = o, o += up(sizeof(double) * M);
    L.T = o, o += up(sizeof(double) * M);
    L.N = o, o += up(sizeof(double) * M);
    L.m2_0 = o, o += up(sizeof(double) * M);
    L.lam = o, o += up(sizeof(double) * M);
    L.bytes = o;
  }
  FS_ALLOC(fs->d_slots, fs->sl.bytes);
  if (hipHostMalloc((void **)&fs->h_slots, fs->sl.bytes, hipHostMallocDefault) != hipSuccess) {
    fs->h_slots = nullptr;
    set_error("flow server: pinned allocation of %zu bytes failed", fs->sl.bytes);
    dfrg_fs_destroy(fs);
    return DFRG_ERR_HIP;
  }
  std::memset(fs->h_slots, 0, fs->sl.bytes);
  d.n_runs = (int *)(fs->d_slots + fs->sl.n_runs);
  d.status = (int *)(fs->d_slots + fs->sl.status);
  d.bad = (int *)(fs->d_slots + fs->sl.bad);
  d.h = (double *)(fs->d_slots + fs->sl.h);
  d.T = (double *)(fs->d_slots + fs->sl.T);
  d.N = (double *)(fs->d_slots + fs->sl.N);
  d.m2_0 = (double *)(fs->d_slots + fs->sl.m2_0);
  d.lam = (double *)(fs->d_slots + fs->sl.lam);
  FS_ALLOC(d.runs, runs_bytes);
  FS_ALLOC(d.base, sizeof(int));
  FS_ALLOC(fs->d_nodes, sizeof(double) * nodes.size());
  FS_ALLOC(d.pool, sizeof(double) * fs->pool_doubles);
#undef FS_ALLOC
  d.A = fs->d_nodes;
  d.D = fs->d_nodes + nx;
  d.kt = fs->d_nodes + 2 * nx;
  if (hipMemcpy(fs->d_nodes, nodes.data(), sizeof(double) * nodes.size(), hipMemcpyHostToDevice) != hipSuccess) {
    set_error("flow server: node table upload failed");
    dfrg_fs_destroy(fs);
    return DFRG_ERR_HIP;
  }
  if (hipStreamCreateWithFlags(&fs->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&fs->ev0) != hipSuccess || hipEventCreate(&fs->ev1) != hipSuccess) {
    set_error("flow server: stream creation failed");
    dfrg_fs_destroy(fs);
    return DFRG_ERR_HIP;
  }
  DFRG_HIP(hipDeviceSynchronize());
  *out = fs;
  return DFRG_OK;
}

int dfrg_fs_destroy(dfrg_fs *fs)
{
  if (!fs) return DFRG_OK;
  hipDeviceSynchronize();
  if (fs->graph_exec) hipGraphExecDestroy(fs->graph_exec);
  if (fs->graph) hipGraphDestroy(fs->graph);
  if (fs->ev0) hipEventDestroy(fs->ev0);
  if (fs->ev1) hipEventDestroy(fs->ev1);
  if (fs->stream) hipStreamDestroy(fs->stream);
  FlowDev &d = fs->dev;
  for (void *p : {(void *)fs->d_slots, (void *)d.runs, (void *)d.base, (void *)d.pool, (void *)fs->d_nodes})
    if (p) hipFree(p);
  for (void *p : {(void *)fs->h_slots, (void *)fs->h_runs, (void *)fs->h_u})
    if (p) hipHostFree(p);
  delete fs;
  return DFRG_OK;
}

int dfrg_fs_submit(dfrg_fs *fs, const dfrg_flow_desc *desc, long long *id)
{
  DFRG_CHECK_ARG(fs && desc && id, "NULL argument");
  DFRG_CHECK_ARG(desc->n_grid >= 4 && desc->n_grid <= 4096, "n_grid %d out of range [4,4096]", desc->n_grid);
  DFRG_CHECK_ARG(desc->rho_max > 0 && desc->T > 0 && desc->N >= 1, "invalid flow parameters");
  DFRG_CHECK_ARG(std::isfinite(desc->m2_0) && std::isfinite(desc->lambda) && std::isfinite(desc->mu),
                 "invalid flow parameters");
  DFRG_CHECK_ARG(dfrg_fs_flow_bytes(desc->n_grid) <= fs->budget, "flow of %d points exceeds the HBM budget",
                 desc->n_grid);
  FlowRecord r;
  r.desc = *desc;
  r.t_submit = now_s();
  const long long fid = fs->next_id++;
  fs->flows.emplace(fid, std::move(r));
  fs->queue.push_back(fid);
  *id = fid;
  return DFRG_OK;
}

int dfrg_fs_progress(dfrg_fs *fs, int *resident, int *queued, int *done)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(fs != nullptr, "NULL server");
  DFRG_TRY(fs_admit(fs));
  if (resident_count(fs) > 0) {
    DFRG_TRY(fs_capture(fs));
    // grid-point steps this launch advances: each resident flow takes the steps it has left, up to
    // steps_per_launch
    for (long long id : fs->slot_flow)
      if (id >= 0) {
        const FlowRecord &r = fs->flows[id];
        const int left = fs->n_steps - (fs->base - r.admit);
        fs->point_steps += (long long)r.desc.n_grid * std::max(0, std::min(left, fs->steps_per_launch));
      }
    // pinned staging for the final u of the flows that take their last step in this launch (no
    // copy is in flight here: the previous launch ended with a synchronisation)
    size_t need = 0;
    for (long long id : fs->slot_flow)
      if (id >= 0) {
        const FlowRecord &r = fs->flows[id];
        if (fs->base + fs->steps_per_launch - r.admit >= fs->n_steps) need += r.desc.n_grid;
      }
    if (need > fs->h_u_cap) {
      const size_t cap = std::max(need, 2 * fs->h_u_cap);
      if (fs->h_u) (void)hipHostFree(fs->h_u);
      fs->h_u = nullptr;
      fs->h_u_cap = 0;
      DFRG_HIP(hipHostMalloc((void **)&fs->h_u, sizeof(double) * cap, hipHostMallocDefault));
      fs->h_u_cap = cap;
    }
    const double t0 = now_s();
    // the slot block and run table go up asynchronously from their pinned staging
    if (fs->slots_dirty) {
      DFRG_HIP(hipMemcpyAsync(fs->d_slots, fs->h_slots, fs->sl.bytes, hipMemcpyHostToDevice, fs->stream));
      fs->slots_dirty = false;
    }
    if (fs->runs_upload > 0)
      DFRG_HIP(hipMemcpyAsync(fs->dev.runs, fs->h_runs, sizeof(FlowRun) * fs->runs_upload, hipMemcpyHostToDevice,
                              fs->stream));
    fs->runs_upload = 0;
    DFRG_HIP(hipEventRecord(fs->ev0, fs->stream));
    DFRG_HIP(hipGraphLaunch(fs->graph_exec, fs->stream));
    DFRG_HIP(hipEventRecord(fs->ev1, fs->stream));
    // behind the graph: the non-finite flags into the mirror, and the finishers' final u
    DFRG_HIP(hipMemcpyAsync(fs->h_slots + fs->sl.bad, fs->dev.bad, sizeof(int) * fs->max_resident,
                            hipMemcpyDeviceToHost, fs->stream));
    fs->u_staged.clear();
    size_t o = 0;
    for (int sl = 0; sl < fs->max_resident; ++sl) {
      const long long id = fs->slot_flow[sl];
      if (id < 0) continue;
      const FlowRecord &r = fs->flows[id];
      if (fs->base + fs->steps_per_launch - r.admit < fs->n_steps) continue;
      const int n = r.desc.n_grid; // u after the last step sits in the buffer of n_steps' parity
      DFRG_HIP(hipMemcpyAsync(fs->h_u + o, fs->dev.pool + r.off + (fs->n_steps & 1) * n, sizeof(double) * n,
                              hipMemcpyDeviceToHost, fs->stream));
      fs->u_staged.push_back({sl, o});
      o += n;
    }
    DFRG_HIP(hipStreamSynchronize(fs->stream));
    fs->base += fs->steps_per_launch;
    fs->busy_s += now_s() - t0;
    float ms = 0.f;
    DFRG_HIP(hipEventElapsedTime(&ms, fs->ev0, fs->ev1));
    fs->device_s += 1e-3 * ms;
    ++fs->launches;
    DFRG_TRY(fs_retire(fs));
    DFRG_TRY(fs_admit(fs));
  }
  if (resident) *resident = resident_count(fs);
  if (queued) *queued = (int)fs->queue.size();
  if (done) *done = fs->finished;
  return DFRG_OK;
}

int dfrg_fs_poll(dfrg_fs *fs, long long id, int *state)
{
  DFRG_CHECK_ARG(fs && state, "NULL argument");
  auto it = fs->flows.find(id);
  DFRG_CHECK_ARG(it != fs->flows.end(), "unknown flow id %lld", id);
  *state = it->second.state;
  return DFRG_OK;
}

int dfrg_fs_result(dfrg_fs *fs, long long id, double *u_out, int n, double *latency_s)
{
  DFRG_CHECK_ARG(fs && u_out, "NULL argument");
  auto it = fs->flows.find(id);
  DFRG_CHECK_ARG(it != fs->flows.end(), "unknown flow id %lld", id);
  FlowRecord &r = it->second;
  DFRG_CHECK_ARG(r.state == 2 || r.state == 3, "flow %lld has not finished", id);
  DFRG_CHECK_ARG(n >= r.desc.n_grid, "output holds %d < %d points", n, r.desc.n_grid);
  std::copy(r.u.begin(), r.u.end(), u_out);
  if (latency_s) *latency_s = r.t_done - r.t_submit;
  const int failed = r.state == 3;
  fs->flows.erase(it);
  if (failed) {
    set_error("flow %lld became non-finite", id);
    return DFRG_ERR_RUNTIME;
  }
  return DFRG_OK;
}

int dfrg_fs_stats(dfrg_fs *fs, int *finished, double *p50_s, double *p99_s, double *busy_s)
{
  DFRG_CHECK_ARG(fs != nullptr, "NULL server");
  std::vector<double> l = fs->latencies;
  std::sort(l.begin(), l.end());
  auto pct = [&](double p) {
    if (l.empty()) return 0.0;
    const double pos = p * (l.size() - 1);
    const size_t i0 = (size_t)pos;
    const size_t i1 = std::min(i0 + 1, l.size() - 1);
    return l[i0] + (pos - i0) * (l[i1] - l[i0]);
  };
  if (finished) *finished = fs->finished;
  if (p50_s) *p50_s = pct(0.50);
  if (p99_s) *p99_s = pct(0.99);
  if (busy_s) *busy_s = fs->busy_s;
  return DFRG_OK;
}

} // extern "C"
