// step.hip -- on-device leapfrog timestep around the tree solve.
//
// The frontends push particles on the host between solves (pepc-mini
// module_helper.f90:205-222 kick-drift, pepc-benchmark module_helper.f90:
// 257-340).  Here the state stays in HBM and one step is
//   kick_drift:    v += dt/2 a;  x += dt v               (state order)
//   grow, traverse at the new positions                  (sorted order)
//   reorder_kick:  state := sorted order (perm), a = q E / m, v += dt/2 a
//   commit:        gathered v, m, label back into the state arrays
//   energy:        per system 1/2 sum q phi and 1/2 sum m v^2
// (kick-drift-kick leapfrog).  Every step ends with the particles in key
// order, as pepc_grow_tree leaves them (module_domains.f90:204-234); label
// keeps each particle's original index.  All launches are plain kernels on
// one stream, so the step can be captured into a graph and replayed.
#include "kernels.h"

namespace pepc {

namespace {

constexpr int kT = 256;

__global__ void kick_drift_kernel(int64_t n, double dt, double* __restrict__ x,
                                  double* __restrict__ y, double* __restrict__ z,
                                  double* __restrict__ vx, double* __restrict__ vy,
                                  double* __restrict__ vz, const double* __restrict__ ax,
                                  const double* __restrict__ ay, const double* __restrict__ az) {
#pragma clang fp contract(off)
  const double h = 0.5 * dt;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double ux = vx[i] + h * ax[i], uy = vy[i] + h * ay[i], uz = vz[i] + h * az[i];
    vx[i] = ux; vy[i] = uy; vz[i] = uz;
    x[i] = x[i] + dt * ux;
    y[i] = y[i] + dt * uy;
    z[i] = z[i] + dt * uz;
  }
}

// Sorted particle i was state particle perm[i] before the grow.  The force on
// a particle is q E (E already carries the interaction sign).
__global__ void reorder_kick_kernel(int64_t n, double dt, const uint32_t* __restrict__ perm,
                                    const double* __restrict__ xs, const double* __restrict__ ys,
                                    const double* __restrict__ zs, const double* __restrict__ qs,
                                    const double* __restrict__ ex, const double* __restrict__ ey,
                                    const double* __restrict__ ez, StepState st) {
#pragma clang fp contract(off)
  const double h = 0.5 * dt;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t p = perm[i];
    const double m = st.m[p];
    const double q = qs[i];
    const double ax = q * ex[i] / m, ay = q * ey[i] / m, az = q * ez[i] / m;
    st.x[i] = xs[i]; st.y[i] = ys[i]; st.z[i] = zs[i]; st.q[i] = q;
    st.ax[i] = ax; st.ay[i] = ay; st.az[i] = az;
    st.tvx[i] = st.vx[p] + h * ax;
    st.tvy[i] = st.vy[p] + h * ay;
    st.tvz[i] = st.vz[p] + h * az;
    st.tm[i] = m;
    st.tlabel[i] = st.label[p];
  }
}

__global__ void commit_kernel(int64_t n, StepState st) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    st.vx[i] = st.tvx[i]; st.vy[i] = st.tvy[i]; st.vz[i] = st.tvz[i];
    st.m[i] = st.tm[i];
    st.label[i] = st.tlabel[i];
  }
}

// Energies in two deterministic passes (fixed summation order, so graph
// replays and eager steps agree bit for bit): one block per chunk of at most
// kEnergyChunk particles of one system writes its partial sums; then one
// block per system adds its chunks' partials (thread t takes chunks t,
// t + kT, ... in order, then a fixed-shape wave and block reduction) into
// record rec = *counter (records past cap are dropped).
__global__ __launch_bounds__(kT) void energy_chunk_kernel(const int64_t* __restrict__ cbeg,
                                                         const double* __restrict__ pot, StepState st,
                                                         double* __restrict__ partial) {
  const int c = blockIdx.x;
  const int64_t b = cbeg[c], e = cbeg[c + 1];
  double ep = 0.0, ek = 0.0;
  for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
    ep += st.q[i] * pot[i];
    ek += st.m[i] * (st.vx[i] * st.vx[i] + st.vy[i] * st.vy[i] + st.vz[i] * st.vz[i]);
  }
  __shared__ double red[2][kT / kWave];
  ep = wave_sum(ep);
  ek = wave_sum(ek);
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) { red[0][w] = ep; red[1][w] = ek; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, d = 0.0;
    for (int k = 0; k < kT / kWave; ++k) { a += red[0][k]; d += red[1][k]; }
    partial[2 * c] = a;
    partial[2 * c + 1] = d;
  }
}
__global__ __launch_bounds__(kT) void energy_system_kernel(const int* __restrict__ sys_chunk,
                                                          const double* __restrict__ partial,
                                                          double* __restrict__ hist, int nsys,
                                                          int64_t cap,
                                                          const uint32_t* __restrict__ counter) {
  const int sid = blockIdx.x;
  const int64_t rec = *counter;
  double a = 0.0, d = 0.0;
  for (int c = sys_chunk[sid] + (int)threadIdx.x; c < sys_chunk[sid + 1]; c += kT) {
    a += partial[2 * c];
    d += partial[2 * c + 1];
  }
  __shared__ double red[2][kT / kWave];
  a = wave_sum(a);
  d = wave_sum(d);
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) { red[0][w] = a; red[1][w] = d; }
  __syncthreads();
  if (threadIdx.x == 0 && rec < cap) {
    double sa = 0.0, sd = 0.0;
    for (int k = 0; k < kT / kWave; ++k) { sa += red[0][k]; sd += red[1][k]; }
    hist[(rec * nsys + sid) * 2 + 0] = 0.5 * sa;
    hist[(rec * nsys + sid) * 2 + 1] = 0.5 * sd;
  }
}
__global__ void counter_inc_kernel(uint32_t* counter) {
  if (threadIdx.x == 0) *counter += 1u;
}

__global__ void counter_set_kernel(uint32_t* counter, uint32_t v) {
  if (threadIdx.x == 0) *counter = v;
}

unsigned grid_for(int64_t n) {
  int64_t b = (n + kT - 1) / kT;
  if (b < 1) b = 1;
  if (b > 8192) b = 8192;
  return (unsigned)b;
}

}  // namespace

void launch_kick_drift(hipStream_t s, int64_t n, double dt, const StepState& st) {
  kick_drift_kernel<<<grid_for(n), kT, 0, s>>>(n, dt, st.x, st.y, st.z, st.vx, st.vy, st.vz, st.ax,
                                               st.ay, st.az);
}

void launch_reorder_kick(hipStream_t s, int64_t n, double dt, const uint32_t* perm, const double* xs,
                         const double* ys, const double* zs, const double* qs, const double* ex,
                         const double* ey, const double* ez, const StepState& st) {
  reorder_kick_kernel<<<grid_for(n), kT, 0, s>>>(n, dt, perm, xs, ys, zs, qs, ex, ey, ez, st);
  commit_kernel<<<grid_for(n), kT, 0, s>>>(n, st);
}

void launch_energy(hipStream_t s, int nsys, int nchunk, const int64_t* cbeg, const int* sys_chunk,
                   const double* pot, const StepState& st, double* partial, double* hist,
                   int64_t cap, uint32_t* counter) {
  if (nchunk > 0) energy_chunk_kernel<<<nchunk, kT, 0, s>>>(cbeg, pot, st, partial);
  if (nsys > 0) energy_system_kernel<<<nsys, kT, 0, s>>>(sys_chunk, partial, hist, nsys, cap, counter);
  counter_inc_kernel<<<1, 64, 0, s>>>(counter);
}

void launch_counter_set(hipStream_t s, uint32_t* counter, uint32_t v) {
  counter_set_kernel<<<1, 64, 0, s>>>(counter, v);
}

}  // namespace pepc
