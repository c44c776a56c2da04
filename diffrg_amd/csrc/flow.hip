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
// Device layout: every resident flow owns 4 n doubles in one HBM pool (u, RK accumulator, two
// stage buffers) and a slot in SoA descriptor arrays.  A point table lists (slot, i) of every
// resident grid point; one wave evaluates four points at a time (one Gauss-Legendre node per lane,
// the eight node sums reduced together across lanes) over a run of consecutive points, so all
// resident flows advance together in each launch.  One RK step is four launches, one per stage;
// steps_per_launch steps are captured into one hipGraph.  A flow's step index is
// the launch's base step (written by the host before each graph launch) plus the step within the
// graph minus the base step at its admission, so flows admitted at different times share the
// graph and a finished flow simply stops updating.  The graph's last node advances the device copy
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
#include <cmath>
#include <deque>
#include <map>
#include <vector>

using namespace dfrg;

namespace
{
  constexpr int kFlowBlock = 256;
  constexpr int kFlowGrid = 1024; // waves = kFlowGrid * kFlowBlock / 64, grid-stride over chunks
  constexpr double kPi = 3.14159265358979323846;

  // One resident grid point: with its flow's slot, state offset, size and admission step, a
  // lane's first load gives the addresses of everything else it reads, so the rest of its data
  // (state values, per-flow constants, k) comes in one second round of loads.
  struct FlowPoint {
    int slot, i, off, n, admit, pad;
  };

  struct FlowDev {
    int *status, *bad;    // per slot; status 0 idle, 1 running (host-owned)
    double *h, *T, *N;    // per slot: grid spacing, temperature, O(N)
    FlowPoint *points;    // every resident grid point
    int *n_points;        // device copy of the point count
    int *base;                           // global step index of the first step of this launch
    double *pool;                        // state: u | acc | y0 | y1 per flow
    const double *A, *D;                 // node factors (header comment)
    const double *kt;                    // k at (step, stage point): 3 per step, t, t + dt/2, t + dt
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

  __global__ __launch_bounds__(kFlowBlock) void flow_stage_kernel(FlowDev F, int stage, int local)
  {
    __shared__ double s_exp2[64]; // 2^(j/64)
    for (int j = threadIdx.x; j < 64; j += kFlowBlock) s_exp2[j] = exp2(j * (1.0 / 64.0));
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave_id = (blockIdx.x * kFlowBlock + threadIdx.x) >> 6;
    const int n_waves = (gridDim.x * kFlowBlock) >> 6;
    const int P = *F.n_points;
    const int base = *F.base + local;
    // A wave takes a chunk of consecutive points, sized so that all waves get work: one point each
    // while points are fewer than waves, up to 64 when they are many.  Lane l gathers the data of
    // point w0 + l (its record, then neighbours in the stage input, its own state and
    // accumulator, the flow's constants and k) in two rounds of loads; the wave then evaluates
    // the points' right-hand sides with the point's values broadcast from its lane (lanes over
    // Gauss-Legendre nodes), and lane l applies the RK stage update of its point.
    const int chunk = min(64, max(1, (P + n_waves - 1) / n_waves));
    for (int w0 = wave_id * chunk; w0 < P; w0 += n_waves * chunk) {
      const int cnt = min(chunk, P - w0);
      int s = 0, i = 0, n = 0, off = 0, active = 0;
      double up = 0.0, upp = 0.0, rho = 0.0, m2pi = 0.0, m2sig = 0.0, T = 1.0, Nn = 0.0, k = 0.0;
      double u_i = 0.0, acc_i = 0.0; // this point's state and RK accumulator
      if (lane < cnt) {
        const FlowPoint pt = F.points[w0 + lane];
        s = pt.slot;
        i = pt.i;
        n = pt.n;
        off = pt.off;
        const int step = base - pt.admit;
        // a listed point's state range is allocated while its flow is resident, so the loads
        // below stay in it; a flow past its last step is left alone
        if (step >= 0 && step < F.n_steps) {
          const double *u = F.pool + off;
          const double *in = stage == 1 ? u : (stage == 3 ? u + 3 * n : u + 2 * n);
          active = F.status[s] == 1;
          const double h = F.h[s];
          T = F.T[s];
          Nn = F.N[s];
          k = F.kt[3 * step + (stage == 1 ? 0 : (stage == 4 ? 2 : 1))];
          u_i = u[i];
          if (stage > 1) acc_i = u[n + i];
          if (i == 0) {
            up = (-3.0 * in[0] + 4.0 * in[1] - in[2]) / (2.0 * h);
            upp = (in[0] - 2.0 * in[1] + in[2]) / (h * h);
          } else if (i == n - 1) {
            up = (3.0 * in[n - 1] - 4.0 * in[n - 2] + in[n - 3]) / (2.0 * h);
            upp = (in[n - 1] - 2.0 * in[n - 2] + in[n - 3]) / (h * h);
          } else {
            up = (in[i + 1] - in[i - 1]) / (2.0 * h);
            upp = (in[i + 1] - 2.0 * in[i] + in[i - 1]) / (h * h);
          }
          rho = i * h;
          m2pi = in[i];
          m2sig = in[i] + 2.0 * rho * up;
        }
      }
      // Points go four at a time: their eight node sums (two threshold integrals each) are
      // reduced together by dev::wave_sum8; integral kind of point pp is vector 2 pp + kind.
      double f_lane = 0.0;
      for (int g0 = 0; g0 < cnt; g0 += 4) {
        double v[8];
#pragma unroll
        for (int pp = 0; pp < 4; ++pp) {
          const int p = g0 + pp;
          double ipi = 0.0, isg = 0.0;
          if (p < cnt && __builtin_amdgcn_readlane(active, p)) { // wave-uniform
            const double pT = dev::lane_value(T, p), pk = dev::lane_value(k, p);
            const double pm2pi = dev::lane_value(m2pi, p), pm2sig = dev::lane_value(m2sig, p);
            // two threshold integrals, lane j evaluates node j (d = 3 radial element)
            const double k2 = pk * pk, inv_T = 1.0 / pT, inv_8T = 0.125 * inv_T;
            for (int j = lane; j < F.nx; j += 64) {
              const double a = F.A[j], dd = F.D[j];
              ipi = fma(dd, bose_S2(fma(k2, a, pm2pi), pT, inv_8T, inv_T, s_exp2), ipi);
              isg = fma(dd, bose_S2(fma(k2, a, pm2sig), pT, inv_8T, inv_T, s_exp2), isg);
            }
          }
          v[2 * pp] = ipi;
          v[2 * pp + 1] = isg;
        }
        const double y = dev::wave_sum8(v);
#pragma unroll
        for (int pp = 0; pp < 4; ++pp) {
          const int p = g0 + pp;
          if (p >= cnt || !__builtin_amdgcn_readlane(active, p)) continue; // wave-uniform
          const double pk = dev::lane_value(k, p);
          const double k5 = pk * pk * (pk * pk) * pk;
          const double ipi = k5 * dev::lane_value(y, dev::sum8_lane(2 * pp));
          const double isg = k5 * dev::lane_value(y, dev::sum8_lane(2 * pp + 1));
          if (lane == p) f_lane = 0.5 * ((Nn - 1.0) * up * ipi + (3.0 * up + 2.0 * rho * upp) * isg);
        }
      }
      if (lane < cnt && active) {
        // only this lane reads or writes u[i] and acc[i] in this launch; the neighbours other
        // lanes read are in the stage's input buffer, which this stage does not write
        double *u = F.pool + off;
        double *acc = u + n, *y0 = u + 2 * n, *y1 = u + 3 * n;
        const double f = f_lane, dt = F.dt;
        if (stage == 1) {
          acc[i] = f;
          y0[i] = u_i + 0.5 * dt * f;
        } else if (stage == 2) {
          acc[i] = acc_i + 2.0 * f;
          y1[i] = u_i + 0.5 * dt * f;
        } else if (stage == 3) {
          acc[i] = acc_i + 2.0 * f;
          y0[i] = u_i + dt * f;
        } else {
          const double un = u_i + dt / 6.0 * (acc_i + f);
          u[i] = un;
          if (!isfinite(un)) atomicOr(&F.bad[s], 1);
        }
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
  bool points_dirty = true;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int base = 0;           // global step index of the next launch's first step (mirrors dev.base)
  double *d_nodes = nullptr; // A | D | kt
  std::vector<double> latencies;
  double busy_s = 0;
  int finished = 0;
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
      for (int stage = 1; stage <= 4; ++stage)
        hipLaunchKernelGGL(flow_stage_kernel, dim3(kFlowGrid), dim3(kFlowBlock), 0, cs, fs->dev, stage, st);
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
    std::vector<int> bad(fs->max_resident);
    DFRG_HIP(hipMemcpy(bad.data(), fs->dev.bad, sizeof(int) * fs->max_resident, hipMemcpyDeviceToHost));
    const double t = now_s();
    for (int s = 0; s < fs->max_resident; ++s) {
      if (fs->slot_flow[s] < 0) continue;
      FlowRecord &r = fs->flows[fs->slot_flow[s]];
      const bool done = fs->base - r.admit >= fs->n_steps;
      if (!done && !bad[s]) continue;
      const int n = r.desc.n_grid;
      r.u.resize(n);
      DFRG_HIP(hipMemcpy(r.u.data(), fs->dev.pool + r.off, sizeof(double) * n, hipMemcpyDeviceToHost));
      r.state = bad[s] ? 3 : 2;
      r.t_done = t;
      if (r.state == 2) {
        fs->latencies.push_back(r.t_done - r.t_submit);
        ++fs->finished;
      }
      fs_free_range(fs, r.off, 4 * n);
      const int zero = 0;
      DFRG_HIP(hipMemcpy(fs->dev.status + s, &zero, sizeof(int), hipMemcpyHostToDevice));
      fs->slot_flow[s] = -1;
      r.slot = -1;
      fs->points_dirty = true;
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
      const int off = fs_alloc_range(fs, 4 * n);
      if (off < 0) break; // does not fit the HBM budget now: wait for retirements (FIFO)
      fs->queue.pop_front();
      std::vector<double> init(4 * n, 0.0);
      const double h = r.desc.rho_max / (n - 1);
      for (int i = 0; i < n; ++i) init[i] = r.desc.m2_0 + r.desc.lambda / 2.0 * (i * h);
      DFRG_HIP(hipMemcpy(fs->dev.pool + off, init.data(), sizeof(double) * 4 * n, hipMemcpyHostToDevice));
      const int one = 1, zero = 0;
      DFRG_HIP(hipMemcpy(fs->dev.bad + slot, &zero, sizeof(int), hipMemcpyHostToDevice));
      DFRG_HIP(hipMemcpy(fs->dev.h + slot, &h, sizeof(double), hipMemcpyHostToDevice));
      DFRG_HIP(hipMemcpy(fs->dev.T + slot, &r.desc.T, sizeof(double), hipMemcpyHostToDevice));
      DFRG_HIP(hipMemcpy(fs->dev.N + slot, &r.desc.N, sizeof(double), hipMemcpyHostToDevice));
      DFRG_HIP(hipMemcpy(fs->dev.status + slot, &one, sizeof(int), hipMemcpyHostToDevice));
      fs->slot_flow[slot] = id;
      r.slot = slot;
      r.off = off;
      r.admit = fs->base;
      r.state = 1;
      fs->points_dirty = true;
    }
    if (fs->points_dirty) {
      std::vector<FlowPoint> pts;
      for (int s = 0; s < fs->max_resident; ++s) {
        if (fs->slot_flow[s] < 0) continue;
        const FlowRecord &r = fs->flows[fs->slot_flow[s]];
        const int n = r.desc.n_grid;
        for (int i = 0; i < n; ++i) pts.push_back(FlowPoint{s, i, r.off, n, r.admit, 0});
      }
      const int P = (int)pts.size();
      if (P) DFRG_HIP(hipMemcpy(fs->dev.points, pts.data(), sizeof(FlowPoint) * P, hipMemcpyHostToDevice));
      DFRG_HIP(hipMemcpy(fs->dev.n_points, &P, sizeof(int), hipMemcpyHostToDevice));
      fs->points_dirty = false;
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

size_t dfrg_fs_flow_bytes(int n_grid) { return (size_t)4 * n_grid * sizeof(double); }

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
  const size_t pts_bytes = sizeof(FlowPoint) * std::min(max_points, fs->pool_doubles / 4 + 1);
#define FS_ALLOC(ptr, bytes)                                                                                   \
  do {                                                                                                         \
    if (hipMalloc((void **)&(ptr), (bytes)) != hipSuccess || hipMemset((ptr), 0, (bytes)) != hipSuccess) {     \
      set_error("flow server: device allocation of %zu bytes failed", (size_t)(bytes));                        \
      dfrg_fs_destroy(fs);                                                                                     \
      return DFRG_ERR_HIP;                                                                                     \
    }                                                                                                          \
  } while (0)
  FS_ALLOC(d.status, sizeof(int) * max_resident);
  FS_ALLOC(d.bad, sizeof(int) * max_resident);
  FS_ALLOC(d.h, sizeof(double) * max_resident);
  FS_ALLOC(d.T, sizeof(double) * max_resident);
  FS_ALLOC(d.N, sizeof(double) * max_resident);
  FS_ALLOC(d.points, pts_bytes);
  FS_ALLOC(d.n_points, sizeof(int));
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
  if (hipStreamCreateWithFlags(&fs->stream, hipStreamNonBlocking) != hipSuccess) {
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
  if (fs->stream) hipStreamDestroy(fs->stream);
  FlowDev &d = fs->dev;
  for (void *p : {(void *)d.status, (void *)d.bad, (void *)d.h,
                  (void *)d.T, (void *)d.N, (void *)d.points, (void *)d.n_points, (void *)d.base, (void *)d.pool,
                  (void *)fs->d_nodes})
    if (p) hipFree(p);
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
    const double t0 = now_s();
    DFRG_HIP(hipGraphLaunch(fs->graph_exec, fs->stream));
    DFRG_HIP(hipStreamSynchronize(fs->stream));
    fs->base += fs->steps_per_launch;
    fs->busy_s += now_s() - t0;
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
