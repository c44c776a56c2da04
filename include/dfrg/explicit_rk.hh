// dfrg/explicit_rk.hh -- explicit Runge-Kutta stepping of "variables" systems (ODEs in the RG time
// t = ln(k/Lambda) whose right-hand sides are loop integrals), the stepper of the reference's
// TimeStepperRK::run_vars (timestepping/rk.cc:57-134):
//   method 1 forward Euler, 2 third-order Runge-Kutta, 3 SSP third order, 4 classical fourth order
//   (the tableaux of deal.II's TimeStepping::runge_kutta_method of those names);
//   fixed step dt from t_start to t_stop, the last step stretched or shortened onto t_stop as
//   deal.II's DiscreteTime does; the state every output_dt (and at t_start) goes to a callback,
//   and a non-finite state at an output step throws (the device form flags the system instead,
//   checking after every step).
//
// Two forms:
//   run_variables      host vector, rhs(t, y, dydt) on the host (which may call the GPU integrators);
//   rk_ensemble<NV>    many independent systems of NV variables on the device, one system per
//                      thread and the whole time integration inside one launch: the user's
//                      __device__ rhs(t, y, dydt, system) evaluates its integrals in-thread, so no
//                      kernel boundary separates the stages and steps.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <functional>
#include <stdexcept>
#include <vector>

namespace dfrg
{
  namespace tmpl
  {
    struct RKTableau {
      int stages;
      double a[4][4]; // a[i][j], j < i
      double b[4];
      double c[4];
    };

    inline RKTableau rk_tableau(const int method)
    {
      switch (method) {
      case 1: return {1, {{0}}, {1.0}, {0.0}};
      case 2: return {3, {{0}, {0.5}, {-1.0, 2.0}}, {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0}, {0.0, 0.5, 1.0}};
      case 3: return {3, {{0}, {1.0}, {0.25, 0.25}}, {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0}, {0.0, 1.0, 0.5}};
      case 4:
        return {4, {{0}, {0.5}, {0.0, 0.5}, {0.0, 0.0, 1.0}}, {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0},
                {0.0, 0.5, 0.5, 1.0}};
      default: throw std::runtime_error("TimeStepperRK::run: method not implemented");
      }
    }

    // The time after t: t + dt, or t_end when t + dt would pass t_end or land within 5% of dt before
    // it (deal.II DiscreteTime's rule: the last step is shortened or stretched onto t_end).
    __host__ __device__ inline double next_time(const double t, const double dt, const double t_end)
    {
      return (t + dt + 0.05 * dt > t_end) ? t_end : t + dt;
    }

    // Every step must advance t by a representable amount and the number of steps stays bounded,
    // so both loops below end (on the device a non-ending loop would hang the GPU).
    inline void check_times(const double t_start, const double t_stop, const double dt)
    {
      if (!(dt > 0) || !std::isfinite(dt)) throw std::invalid_argument("TimeStepperRK::run: dt must be positive");
      if (!std::isfinite(t_start) || !std::isfinite(t_stop))
        throw std::invalid_argument("TimeStepperRK::run: start and stop times must be finite");
      if (t_stop > t_start &&
          ((t_stop - t_start) / dt > 1e7 || t_start + dt == t_start || t_stop - dt == t_stop))
        throw std::invalid_argument("TimeStepperRK::run: dt too small for the time interval (at most 1e7 steps)");
    }

    // Host stepper.  rhs(t, y, dydt); output(t, y) at t_start and every output_each steps, where
    // output_each = floor(output_dt / dt) (at least 1).
    inline void run_variables(std::vector<double> &y, const double t_start, const double t_stop, const double dt,
                              const double output_dt, const int method,
                              const std::function<void(double, const std::vector<double> &, std::vector<double> &)> &rhs,
                              const std::function<void(double, const std::vector<double> &)> &output = nullptr)
    {
      if (y.empty()) throw std::runtime_error("TimeStepperRK::run: y contains no variables, use a different timestepper!");
      check_times(t_start, t_stop, dt);
      const RKTableau tab = rk_tableau(method);
      const size_t n = y.size();
      const unsigned output_each = std::max(1u, (unsigned)(output_dt / dt));
      std::vector<std::vector<double>> k(tab.stages, std::vector<double>(n));
      std::vector<double> Y(n);
      if (output) output(t_start, y);
      double t = t_start;
      unsigned step = 0;
      while (t < t_stop) {
        const double t_next = next_time(t, dt, t_stop), h = t_next - t;
        for (int i = 0; i < tab.stages; ++i) {
          for (size_t v = 0; v < n; ++v) {
            double s = y[v];
            for (int j = 0; j < i; ++j) s += h * tab.a[i][j] * k[j][v];
            Y[v] = s;
          }
          rhs(t + tab.c[i] * h, Y, k[i]);
        }
        for (size_t v = 0; v < n; ++v) {
          double s = 0.0;
          for (int i = 0; i < tab.stages; ++i) s += tab.b[i] * k[i][v];
          y[v] += h * s;
        }
        t = t_next;
        ++step;
        if (step % output_each == 0) {
          double norm = 0.0;
          for (double v : y) norm += v * v;
          if (!std::isfinite(norm)) throw std::runtime_error("TimeStepperRK::run: NaNs in solution");
          if (output) output(t, y);
        }
      }
    }

    // One system per thread.  y: n_systems x NV (row per system), updated in place.  Snapshots:
    // if out != nullptr, the state at t_start and after every output_each steps is written to
    // out[(o * n_systems + s) * NV + v], o < n_out.  status[s] = 0, or 1 if the state stopped
    // being finite (that system stops).
    template <int NV, int STAGES, class RHS>
    __global__ __launch_bounds__(256) void rk_ensemble_kernel(RHS rhs, RKTableau tab, double *y, int n_systems,
                                                              double t_start, double t_stop, double dt,
                                                              int output_each, double *out, int n_out, int *status)
    {
      const int s = blockIdx.x * blockDim.x + threadIdx.x;
      if (s >= n_systems) return;
      double ys[NV], Y[NV], k[STAGES][NV];
      for (int v = 0; v < NV; ++v) ys[v] = y[(size_t)s * NV + v];
      int o = 0;
      if (out && o < n_out) {
        for (int v = 0; v < NV; ++v) out[((size_t)o * n_systems + s) * NV + v] = ys[v];
        ++o;
      }
      int bad = 0;
      double t = t_start;
      int step = 0;
      while (t < t_stop) {
        const double t_next = next_time(t, dt, t_stop), h = t_next - t;
#pragma unroll
        for (int i = 0; i < STAGES; ++i) {
          for (int v = 0; v < NV; ++v) {
            double acc = ys[v];
#pragma unroll
            for (int j = 0; j < i; ++j) acc += h * tab.a[i][j] * k[j][v];
            Y[v] = acc;
          }
          rhs(t + tab.c[i] * h, Y, k[i], s);
        }
        double norm = 0.0;
        for (int v = 0; v < NV; ++v) {
          double acc = 0.0;
#pragma unroll
          for (int i = 0; i < STAGES; ++i) acc += tab.b[i] * k[i][v];
          ys[v] += h * acc;
          norm += ys[v] * ys[v];
        }
        t = t_next;
        ++step;
        if (!isfinite(norm)) {
          bad = 1;
          break;
        }
        if (out && step % output_each == 0 && o < n_out) {
          for (int v = 0; v < NV; ++v) out[((size_t)o * n_systems + s) * NV + v] = ys[v];
          ++o;
        }
      }
      for (int v = 0; v < NV; ++v) y[(size_t)s * NV + v] = ys[v];
      if (status) status[s] = bad;
    }

    // Number of snapshots rk_ensemble writes for these times (t_start included).
    inline int rk_snapshots(const double t_start, const double t_stop, const double dt, const double output_dt)
    {
      check_times(t_start, t_stop, dt);
      const int output_each = std::max(1, (int)(output_dt / dt));
      int steps = 0;
      for (double t = t_start; t < t_stop; ++steps) t = next_time(t, dt, t_stop);
      return 1 + steps / output_each;
    }

    // Launch the ensemble on `stream` (device pointers; y updated in place).
    template <int NV, class RHS>
    void rk_ensemble(const RHS &rhs, const int method, double *d_y, const int n_systems, const double t_start,
                     const double t_stop, const double dt, const double output_dt, double *d_out, const int n_out,
                     int *d_status, hipStream_t stream)
    {
      static_assert(NV >= 1 && NV <= 64, "rk_ensemble: 1 to 64 variables per system");
      check_times(t_start, t_stop, dt);
      if (n_systems <= 0) return;
      if (n_out > 0 && d_out == nullptr) throw std::invalid_argument("rk_ensemble: snapshots requested without a buffer");
      const RKTableau tab = rk_tableau(method);
      const int output_each = std::max(1, (int)(output_dt / dt));
      const dim3 grid((n_systems + 255) / 256), block(256);
      switch (tab.stages) {
      case 1:
        hipLaunchKernelGGL((rk_ensemble_kernel<NV, 1, RHS>), grid, block, 0, stream, rhs, tab, d_y, n_systems,
                           t_start, t_stop, dt, output_each, d_out, n_out, d_status);
        break;
      case 3:
        hipLaunchKernelGGL((rk_ensemble_kernel<NV, 3, RHS>), grid, block, 0, stream, rhs, tab, d_y, n_systems,
                           t_start, t_stop, dt, output_each, d_out, n_out, d_status);
        break;
      default:
        hipLaunchKernelGGL((rk_ensemble_kernel<NV, 4, RHS>), grid, block, 0, stream, rhs, tab, d_y, n_systems,
                           t_start, t_stop, dt, output_each, d_out, n_out, d_status);
        break;
      }
      if (hipGetLastError() != hipSuccess) throw std::runtime_error("rk_ensemble: launch failed");
    }
  } // namespace tmpl
} // namespace dfrg
