// scmf_engine.hip — HIP kernels for gfx950 and the C ABI of include/scmf.h.
//
// One system lives on one device.  Device layout:
//   x[], y[], z[]  fp32, n_beads each; chain c owns beads
//                  [bead_base[p] + (c - chain_base[p]) N_p, ... + N_p)
//   counts         uint16 pairs packed in uint32 words, or uint32, [type][z][y][x]
//   omega          fp32 [type][z][y][x]
// Kernels:
//   k_init_coords  one lane per chain: start point + Gaussian walk
//   k_sweep        one lane per chain: N sequential smart-MC / trial moves
//   k_zero         clear the counts (a kernel, not a memset, so it can sit in a graph)
//   k_grid         one lane per bead: NGP cell, atomic increment
//   k_omega        one lane per cell: all T omega values in fp64, stored fp32
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/scmf.h"
#include "scmf_device.h"
#include "scmf_host.h"

#pragma clang fp contract(off)

using namespace scmf;

// ============================================================== kernels
namespace {

constexpr int BLOCK = 256;
constexpr int MAX_INIT_N = 4096;  // occupancy bitmask of the initial walk

// initialize_coord, device_function.cu:140-207, one lane per chain.
__global__ void __launch_bounds__(BLOCK) k_init_coords(KParams kp, float* __restrict__ x,
                                                       float* __restrict__ y,
                                                       float* __restrict__ z) {
  const int64_t chain = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (chain >= kp.n_chains) return;
  const int p = arch_of_chain(kp, chain);
  const int N = kp.len[p];
  const int64_t pi = kp.bead_base[p] + (chain - kp.chain_base[p]) * N;
  uint32_t occ[MAX_INIT_N / 32];
  for (int w = 0; w < (N + 31) / 32; ++w) occ[w] = 0;
  const u32x4 w0 = stream_block(kp.seed, kp.system_id, (uint32_t)chain, INIT_SWEEP, 0, SLOT_INIT);
  x[pi] = (float)((double)u01_24f(w0.v[0]) * kp.Lx);
  y[pi] = (float)((double)u01_24f(w0.v[1]) * kp.Ly);
  z[pi] = (float)((double)u01_24f(w0.v[2]) * kp.Lz);
  occ[0] = 1u;
  const double cste = 1.0 / sqrt(3.0 * ((double)N - 1.0));
  const int mo = kp.mono_off[p];
  for (int i = 0; i < N; ++i) {
    for (int k = kp.nbr_start[mo + i]; k < kp.nbr_start[mo + i + 1]; ++k) {
      const int j = kp.nbr[k];
      if (((occ[j >> 5] >> (j & 31)) & 1u) == 0) {
        const u32x4 w =
            stream_block(kp.seed, kp.system_id, (uint32_t)chain, INIT_SWEEP, (uint32_t)(1 + j), SLOT_INIT);
        double n0, n1, n2;
        normal3(w, &n0, &n1, &n2);
        x[pi + j] = (float)(n0 * cste + (double)x[pi + i]);
        y[pi + j] = (float)(n1 * cste + (double)y[pi + i]);
        z[pi + j] = (float)(n2 * cste + (double)z[pi + i]);
        occ[j >> 5] |= 1u << (j & 31);
      }
    }
  }
}

// add_bond_forces_arr, device_function.cu:589-634.
__device__ __forceinline__ void bond_forces(const KParams& kp, int b, int64_t poly_index,
                                            const float* x, const float* y, const float* z,
                                            float px, float py, float pz, float* fx, float* fy,
                                            float* fz) {
  const double harm = kp.harm;
  double v1x = 0.0, v1y = 0.0, v1z = 0.0;
  const int k0 = kp.nbr_start[b], k1 = kp.nbr_start[b + 1];
  for (int k = k0; k < k1; ++k) {
    const int64_t ni = poly_index + kp.nbr[k];
    v1x += (double)(x[ni] - px) * 2.0 * harm;
    v1y += (double)(y[ni] - py) * 2.0 * harm;
    v1z += (double)(z[ni] - pz) * 2.0 * harm;
  }
  *fx = (float)((double)*fx + v1x);
  *fy = (float)((double)*fy + v1y);
  *fz = (float)((double)*fz + v1z);
}

// mc_polymer_move_arr, device_function.cu:718-799, one lane per chain, with
// trial_move_smc_arr (:637-717), calc_delta_energy_arr (:562-588),
// calc_delta_bonded_energy_arr (:509-561) and som_accept (:444-454) inlined.
__global__ void __launch_bounds__(BLOCK) k_sweep(KParams kp, float* __restrict__ x,
                                                 float* __restrict__ y, float* __restrict__ z,
                                                 const float* __restrict__ omega, uint32_t sweep,
                                                 unsigned long long* __restrict__ accepts) {
  const int64_t chain = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  uint32_t acc = 0;
  if (chain < kp.n_chains) {
    const int p = arch_of_chain(kp, chain);
    const int length = kp.len[p];
    const int64_t poly_index = kp.bead_base[p] + (chain - kp.chain_base[p]) * length;
    const int mo = kp.mono_off[p];
    for (int i = 0; i < length; i++) {
      const u32x4 pick =
          stream_block(kp.seed, kp.system_id, (uint32_t)chain, sweep, (uint32_t)i, SLOT_PICK);
      int ibead = (int)(u01_53(pick.v[0], pick.v[1]) * length);
      if (ibead == length) ibead = length - 1;
      const uint32_t iwtype = (uint32_t)kp.monotype[mo + ibead];
      const int64_t index = poly_index + ibead;
      const int b = mo + ibead;
      const u32x4 disp =
          stream_block(kp.seed, kp.system_id, (uint32_t)chain, sweep, (uint32_t)i, SLOT_DISP);
      const float ox = x[index], oy = y[index], oz = z[index];
      float dx, dy, dz;
      double smc_deltaE = 0.0;
      if (kp.move_type == SCMF_MOVE_SMART) {
        float px = ox, py = oy, pz = oz;
        const float A = kp.A[iwtype];
        const float R = kp.R[iwtype];
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        bond_forces(kp, b, poly_index, x, y, z, px, py, pz, &fx, &fy, &fz);
        double n0, n1, n2;
        normal3(disp, &n0, &n1, &n2);
        const float rx = (float)n0, ry = (float)n1, rz = (float)n2;
        dx = A * fx + rx * R;
        dy = A * fy + ry * R;
        dz = A * fz + rz * R;
        px += dx;
        py += dy;
        pz += dz;
        float nfx = 0.0f, nfy = 0.0f, nfz = 0.0f;
        bond_forces(kp, b, poly_index, x, y, z, px, py, pz, &nfx, &nfy, &nfz);
        smc_deltaE = 0.0;
        smc_deltaE += 0.5 * ((nfx + fx) * dx + (nfy + fy) * dy + (nfz + fz) * dz);
        smc_deltaE += 0.25 * A *
                      ((nfx * nfx) + (nfy * nfy) + (nfz * nfz) - (fx * fx) - (fy * fy) - (fz * fz));
      } else {
        const float delta = kp.trial_delta;
        dx = delta * (2.0f * u01_24f_closed0(disp.v[0]) - 1.0f);
        dy = delta * (2.0f * u01_24f_closed0(disp.v[1]) - 1.0f);
        dz = delta * (2.0f * u01_24f_closed0(disp.v[2]) - 1.0f);
      }
      const float newx = ox + dx, newy = oy + dy, newz = oz + dz;
      const double xold = ox, yold = oy, zold = oz;
      const uint32_t cell_old = bin_cell(kp, xold, yold, zold, iwtype);
      const double energy_old = omega[cell_old];
      const uint32_t cell_new =
          bin_cell(kp, xold + (double)dx, yold + (double)dy, zold + (double)dz, iwtype);
      const double energy_new = omega[cell_new];
      double bonded = 0;
      {
        const double ddx = dx, ddy = dy, ddz = dz;
        const int k0 = kp.nbr_start[b], k1 = kp.nbr_start[b + 1];
        for (int k = k0; k < k1; ++k) {
          const int64_t nb = poly_index + kp.nbr[k];
          const double old_rx = (double)(ox - x[nb]);
          const double new_rx = old_rx + ddx;
          const double old_ry = (double)(oy - y[nb]);
          const double new_ry = old_ry + ddy;
          const double old_rz = (double)(oz - z[nb]);
          const double new_rz = old_rz + ddz;
          const double old_r2 = old_rx * old_rx + old_ry * old_ry + old_rz * old_rz;
          const double new_r2 = new_rx * new_rx + new_ry * new_ry + new_rz * new_rz;
          bonded += kp.harm * (new_r2 - old_r2);
        }
      }
      double energy = energy_new - energy_old;
      energy += bonded;
      double delta_energy = energy;
      delta_energy += smc_deltaE;
      const double p_acc = exp_cw(-1.0 * delta_energy);
      if ((p_acc > 1) || (p_acc > (double)u01_24f(pick.v[2]))) {
        x[index] = newx;
        y[index] = newy;
        z[index] = newz;
        ++acc;
      }
    }
  }
  // Wavefront sum of accepted moves, one atomic per wavefront.
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0 && acc) atomicAdd(accepts, (unsigned long long)acc);
}

__global__ void __launch_bounds__(BLOCK) k_zero(uint32_t* __restrict__ p, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

// coord_to_density, device_function.cu:479-508, one lane per bead.
template <int BITS>
__global__ void __launch_bounds__(BLOCK) k_grid(KParams kp, const float* __restrict__ x,
                                                const float* __restrict__ y,
                                                const float* __restrict__ z,
                                                uint32_t* __restrict__ words,
                                                uint32_t* __restrict__ overflow) {
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t g = (int64_t)blockIdx.x * BLOCK + threadIdx.x; g < kp.n_beads; g += stride) {
    const int p = arch_of_bead(kp, g);
    const int m = (int)((g - kp.bead_base[p]) % kp.len[p]);
    const uint32_t t = (uint32_t)kp.monotype[kp.mono_off[p] + m];
    const uint32_t cell = bin_cell(kp, (double)x[g], (double)y[g], (double)z[g], t);
    if (BITS == 32) {
      atomicAdd(&words[cell], 1u);
    } else {
      const uint32_t sh = (cell & 1u) * 16u;
      const uint32_t old = atomicAdd(&words[cell >> 1], 1u << sh);
      if (((old >> sh) & 0xFFFFu) == 0xFFFFu) atomicOr(overflow, 1u);
    }
  }
}

template <int BITS>
__device__ __forceinline__ uint32_t load_count(const uint32_t* words, uint32_t idx) {
  if (BITS == 32) return words[idx];
  return (words[idx >> 1] >> ((idx & 1u) * 16u)) & 0xFFFFu;
}

// omega_field_update, device_function.cu:373-412: fp64 in the reference's
// operation order, rounded once to fp32.
template <int BITS>
__global__ void __launch_bounds__(BLOCK) k_omega(KParams kp, const uint32_t* __restrict__ words,
                                                 float* __restrict__ omega) {
  const uint32_t C = kp.n_cells;
  const int T = kp.T;
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t c = (int64_t)blockIdx.x * BLOCK + threadIdx.x; c < C; c += stride) {
    double n[MAX_T];
    for (int s = 0; s < T; ++s) n[s] = (double)load_count<BITS>(words, (uint32_t)c + s * C);
    for (int t = 0; t < T; ++t) {
      double w = 0;
      for (int s = 0; s < T; ++s) w += n[s] * kp.fst[t];
      w += -1;
      w = kp.inv_nref * (kp.xn[t + T * t] * w);
      for (int s = 0; s < T; ++s) {
        if (t != s) {
          const double dnorm = -0.5 * kp.xn[t + T * s];
          const double interaction = dnorm * kp.inv_nref * (kp.fst[t] * n[t] - kp.fst[s] * n[s]);
          w += interaction;
        }
      }
      omega[c + (int64_t)t * C] = (float)w;
    }
  }
}

// Mean squared end-to-end distance: per-block partial sums (summed on the
// host in block order, so the result is reproducible).
__global__ void __launch_bounds__(BLOCK) k_re2(KParams kp, const float* __restrict__ x,
                                               const float* __restrict__ y,
                                               const float* __restrict__ z,
                                               double* __restrict__ partial) {
  __shared__ double red[BLOCK];
  const int64_t chain = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  double v = 0;
  if (chain < kp.n_chains) {
    const int p = arch_of_chain(kp, chain);
    const int64_t a = kp.bead_base[p] + (chain - kp.chain_base[p]) * kp.len[p];
    const int64_t e = a + kp.len[p] - 1;
    const double ddx = (double)x[a] - (double)x[e];
    const double ddy = (double)y[a] - (double)y[e];
    const double ddz = (double)z[a] - (double)z[e];
    v = ddx * ddx + ddy * ddy + ddz * ddz;
  }
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(BLOCK) k_planar_to_aos(const float* __restrict__ x,
                                                         const float* __restrict__ y,
                                                         const float* __restrict__ z,
                                                         float* __restrict__ aos, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) {
    aos[3 * i] = x[i];
    aos[3 * i + 1] = y[i];
    aos[3 * i + 2] = z[i];
  }
}

__global__ void __launch_bounds__(BLOCK) k_aos_to_planar(const float* __restrict__ aos,
                                                         float* __restrict__ x,
                                                         float* __restrict__ y,
                                                         float* __restrict__ z, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) {
    x[i] = aos[3 * i];
    y[i] = aos[3 * i + 1];
    z[i] = aos[3 * i + 2];
  }
}

template <int BITS>
__global__ void __launch_bounds__(BLOCK) k_widen(const uint32_t* __restrict__ words,
                                                 uint32_t* __restrict__ out, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride)
    out[i] = load_count<BITS>(words, (uint32_t)i);
}

}  // namespace

// ============================================================== host side
namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(SCMF_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
  } while (0)

enum Kind { K_SWEEP = 0, K_GRID = 1, K_OMEGA = 2, K_KINDS = 3 };

struct Timed {
  int kind;
  hipEvent_t a, b;
};

unsigned grid_for(int64_t n) {
  int64_t g = (n + BLOCK - 1) / BLOCK;
  if (g < 1) g = 1;
  if (g > (1 << 30)) g = 1 << 30;
  return (unsigned)g;
}

unsigned grid_strided(int64_t n) {
  int64_t g = (n + BLOCK - 1) / BLOCK;
  if (g < 1) g = 1;
  if (g > 1024 * 64) g = 1024 * 64;
  return (unsigned)g;
}

}  // namespace

struct scmf_sys {
  int device = 0;
  hipStream_t stream = nullptr;
  HostDesc desc;
  scmf_opts opts{};
  KParams kp{};
  int64_t n_beads = 0, n_chains = 0, n_cells = 0;
  int T = 0;
  float* x = nullptr;
  float* y = nullptr;
  float* z = nullptr;
  uint32_t* counts = nullptr;
  int64_t count_words = 0;
  float* omega = nullptr;
  int* d_monotype = nullptr;
  int* d_nbr_start = nullptr;
  int* d_nbr = nullptr;
  unsigned long long* d_accepts = nullptr;
  uint32_t* d_overflow = nullptr;
  double* d_partial = nullptr;
  float* staging = nullptr;
  int64_t staging_beads = 0;
  uint64_t sweeps = 0;
  uint64_t moves = 0;
  bool timing = false;
  std::vector<Timed> pending;
  std::vector<hipEvent_t> pool;
  double t_sum[K_KINDS] = {0, 0, 0};
  uint64_t t_n[K_KINDS] = {0, 0, 0};

  int ev_begin(int kind, Timed* t) {
    t->kind = kind;
    for (hipEvent_t* e : {&t->a, &t->b}) {
      if (!pool.empty()) {
        *e = pool.back();
        pool.pop_back();
      } else {
        HIP_TRY(hipEventCreate(e));
      }
    }
    HIP_TRY(hipEventRecord(t->a, stream));
    return SCMF_OK;
  }
  int ev_end(Timed* t) {
    HIP_TRY(hipEventRecord(t->b, stream));
    pending.push_back(*t);
    return SCMF_OK;
  }
  int drain_timing() {
    if (pending.empty()) return SCMF_OK;
    HIP_TRY(hipStreamSynchronize(stream));
    for (Timed& t : pending) {
      float ms = 0;
      HIP_TRY(hipEventElapsedTime(&ms, t.a, t.b));
      t_sum[t.kind] += ms;
      t_n[t.kind] += 1;
      pool.push_back(t.a);
      pool.push_back(t.b);
    }
    pending.clear();
    return SCMF_OK;
  }
};

namespace {

int derive(scmf_sys* s) {
  const HostDesc& d = s->desc;
  KParams& kp = s->kp;
  memset(&kp, 0, sizeof kp);
  kp.T = d.T;
  kp.P = (int)d.poly_length.size();
  kp.nx = d.n[0];
  kp.ny = d.n[1];
  kp.nz = d.n[2];
  kp.Lx = d.L[0];
  kp.Ly = d.L[1];
  kp.Lz = d.L[2];
  kp.dx = kp.Lx / kp.nx;
  kp.dy = kp.Ly / kp.ny;
  kp.dz = kp.Lz / kp.nz;
  s->n_cells = (int64_t)kp.nx * kp.ny * kp.nz;
  s->T = d.T;
  if ((int64_t)d.T * s->n_cells >= (int64_t)0xFFFFFFFF)
    return fail(SCMF_E_ARG, "types x cells must stay below 2^32");
  kp.n_cells = (uint32_t)s->n_cells;
  // init_config.cu:480-484 and :662-664.
  const double cste = 1.0 / sqrt(3.0 * ((double)d.nref - 1.0));
  kp.harm = 1.0 / (2.0 * cste * cste);
  kp.inv_nref = 1 / (double)d.nref;
  for (int t = 0; t < d.T; ++t)
    for (int u = 0; u < d.T; ++u) kp.xn[t + d.T * u] = d.xn[t + (size_t)d.T * u];
  int64_t cb = 0, bb = 0;
  int mo = 0;
  for (int p = 0; p < kp.P; ++p) {
    kp.chain_base[p] = cb;
    kp.bead_base[p] = bb;
    kp.mono_off[p] = mo;
    kp.len[p] = d.poly_length[p];
    cb += d.n_polymers[p];
    bb += d.n_polymers[p] * (int64_t)d.poly_length[p];
    mo += d.poly_length[p];
  }
  kp.chain_base[kp.P] = cb;
  kp.bead_base[kp.P] = bb;
  kp.mono_off[kp.P] = mo;
  s->n_chains = cb;
  s->n_beads = bb;
  kp.n_chains = cb;
  kp.n_beads = bb;
  // init_config.cu:784-785 and :804-809.
  for (int t = 0; t < d.T; ++t) {
    kp.fst[t] = (double)s->n_cells / (double)s->n_beads;
    const double A = d.dtNxi / d.nref;
    kp.A[t] = (float)A;
    kp.R[t] = (float)sqrt(A * 2);
  }
  kp.seed = s->opts.seed;
  kp.system_id = s->opts.system_id;
  kp.move_type = s->opts.move_type;
  kp.trial_delta = s->opts.trial_delta;
  return SCMF_OK;
}

int upload_arch(scmf_sys* s) {
  const HostDesc& d = s->desc;
  std::vector<int> nbr_start, nbr;
  size_t coff = 0;
  for (size_t p = 0; p < d.poly_length.size(); ++p) {
    const int N = d.poly_length[p];
    for (int i = 0; i < N; ++i) {
      nbr_start.push_back((int)nbr.size());
      for (int j = 0; j < N; ++j)
        if (d.connection[coff + (size_t)i * N + j]) nbr.push_back(j);
    }
    coff += (size_t)N * N;
  }
  nbr_start.push_back((int)nbr.size());
  if (nbr.empty()) nbr.push_back(0);
  std::vector<int> mono(d.monotype.begin(), d.monotype.end());
  HIP_TRY(hipMalloc(&s->d_monotype, sizeof(int) * mono.size()));
  HIP_TRY(hipMalloc(&s->d_nbr_start, sizeof(int) * nbr_start.size()));
  HIP_TRY(hipMalloc(&s->d_nbr, sizeof(int) * nbr.size()));
  HIP_TRY(hipMemcpy(s->d_monotype, mono.data(), sizeof(int) * mono.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s->d_nbr_start, nbr_start.data(), sizeof(int) * nbr_start.size(),
                    hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s->d_nbr, nbr.data(), sizeof(int) * nbr.size(), hipMemcpyHostToDevice));
  s->kp.monotype = s->d_monotype;
  s->kp.nbr_start = s->d_nbr_start;
  s->kp.nbr = s->d_nbr;
  return SCMF_OK;
}

int ensure_staging(scmf_sys* s, int64_t beads) {
  if (s->staging_beads >= beads) return SCMF_OK;
  if (s->staging) HIP_TRY(hipFree(s->staging));
  s->staging = nullptr;
  s->staging_beads = 0;
  HIP_TRY(hipMalloc(&s->staging, sizeof(float) * 3 * beads));
  s->staging_beads = beads;
  return SCMF_OK;
}

constexpr int64_t STAGE_BEADS = (int64_t)1 << 24;

int free_all(scmf_sys* s) {
  for (void* p : {(void*)s->x, (void*)s->y, (void*)s->z, (void*)s->counts, (void*)s->omega,
                  (void*)s->d_monotype, (void*)s->d_nbr_start, (void*)s->d_nbr,
                  (void*)s->d_accepts, (void*)s->d_overflow, (void*)s->d_partial,
                  (void*)s->staging})
    if (p) (void)hipFree(p);
  for (Timed& t : s->pending) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (hipEvent_t e : s->pool) (void)hipEventDestroy(e);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  return SCMF_OK;
}

int create_impl(scmf_sys** out, const HostDesc& hd, const scmf_opts* opts) {
  if (!out) return fail(SCMF_E_ARG, "out is null");
  *out = nullptr;
  scmf_desc view = hd.view();
  std::string v = validate_desc(&view);
  if (!v.empty()) return fail(SCMF_E_ARG, v);
  if (hd.T > MAX_T) return fail(SCMF_E_ARG, "at most 8 monomer types");
  if ((int)hd.poly_length.size() > MAX_P) return fail(SCMF_E_ARG, "at most 16 polymer types");
  scmf_opts o{};
  if (opts) o = *opts;
  else {
    o.move_type = SCMF_MOVE_SMART;
    o.density_bits = 32;
    o.init_coords = 1;
  }
  if (o.density_bits != 16 && o.density_bits != 32)
    return fail(SCMF_E_ARG, "density_bits must be 16 or 32");
  if (o.move_type != SCMF_MOVE_SMART && o.move_type != SCMF_MOVE_TRIAL)
    return fail(SCMF_E_ARG, "unknown move type");
  if (o.move_type == SCMF_MOVE_TRIAL && !(o.trial_delta > 0))
    return fail(SCMF_E_ARG, "trial moves need trial_delta > 0");
  for (int32_t N : hd.poly_length)
    if (o.init_coords && N > MAX_INIT_N)
      return fail(SCMF_E_ARG, "initial walk supports chains up to 4096 beads");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return fail(SCMF_E_NODEV, "no HIP device available");
  scmf_sys* s = new scmf_sys();
  s->desc = hd;
  s->opts = o;
  int rc = derive(s);
  auto bail = [&](int code) {
    free_all(s);
    delete s;
    return code;
  };
  if (rc) return bail(rc);
  if (hipGetDevice(&s->device) != hipSuccess) return bail(fail(SCMF_E_HIP, "hipGetDevice failed"));
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(SCMF_E_HIP, "hipStreamCreate failed"));
  if ((rc = upload_arch(s))) return bail(rc);
  const int64_t B = s->n_beads;
  const int64_t TC = (int64_t)s->T * s->n_cells;
  s->count_words = o.density_bits == 32 ? TC : (TC + 1) / 2;
  auto alloc = [&](void** p, size_t bytes) -> int {
    if (hipMalloc(p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return fail(SCMF_E_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    }
    return SCMF_OK;
  };
  if ((rc = alloc((void**)&s->x, sizeof(float) * B)) || (rc = alloc((void**)&s->y, sizeof(float) * B)) ||
      (rc = alloc((void**)&s->z, sizeof(float) * B)) ||
      (rc = alloc((void**)&s->counts, sizeof(uint32_t) * s->count_words)) ||
      (rc = alloc((void**)&s->omega, sizeof(float) * TC)) ||
      (rc = alloc((void**)&s->d_accepts, sizeof(unsigned long long))) ||
      (rc = alloc((void**)&s->d_overflow, sizeof(uint32_t))) ||
      (rc = alloc((void**)&s->d_partial, sizeof(double) * grid_for(s->n_chains))))
    return bail(rc);
  if (hipMemsetAsync(s->d_accepts, 0, sizeof(unsigned long long), s->stream) != hipSuccess ||
      hipMemsetAsync(s->d_overflow, 0, sizeof(uint32_t), s->stream) != hipSuccess ||
      hipMemsetAsync(s->omega, 0, sizeof(float) * TC, s->stream) != hipSuccess ||
      hipMemsetAsync(s->counts, 0, sizeof(uint32_t) * s->count_words, s->stream) != hipSuccess)
    return bail(fail(SCMF_E_HIP, "initial clear failed"));
  if (o.init_coords) {
    hipLaunchKernelGGL(k_init_coords, dim3(grid_for(s->n_chains)), dim3(BLOCK), 0, s->stream, s->kp,
                       s->x, s->y, s->z);
    if (hipGetLastError() != hipSuccess) return bail(fail(SCMF_E_HIP, "init launch failed"));
  } else {
    if (hipMemsetAsync(s->x, 0, sizeof(float) * B, s->stream) != hipSuccess ||
        hipMemsetAsync(s->y, 0, sizeof(float) * B, s->stream) != hipSuccess ||
        hipMemsetAsync(s->z, 0, sizeof(float) * B, s->stream) != hipSuccess)
      return bail(fail(SCMF_E_HIP, "position clear failed"));
  }
  if (hipStreamSynchronize(s->stream) != hipSuccess)
    return bail(fail(SCMF_E_HIP, "initialisation failed on the device"));
  *out = s;
  return SCMF_OK;
}

}  // namespace

// ============================================================== C ABI
extern "C" {

const char* scmf_last_error(void) { return g_err.c_str(); }
const char* scmf_version(void) { return "scmf-mi355x 0.1 (gfx950)"; }

int scmf_create(scmf_sys** out, const scmf_desc* desc, const scmf_opts* opts) {
  if (!desc) return fail(SCMF_E_ARG, "desc is null");
  std::string v = validate_desc(desc);
  if (!v.empty()) return fail(SCMF_E_ARG, v);
  HostDesc hd;
  hd.T = desc->n_mono_types;
  hd.xn.assign(desc->xn, desc->xn + (size_t)hd.T * hd.T);
  for (int i = 0; i < 3; ++i) {
    hd.L[i] = desc->L[i];
    hd.n[i] = desc->n[i];
  }
  hd.dtNxi = desc->dtNxi;
  hd.nref = desc->reference_Nbeads;
  hd.poly_length.assign(desc->poly_length, desc->poly_length + desc->n_poly_types);
  hd.n_polymers.assign(desc->n_polymers, desc->n_polymers + desc->n_poly_types);
  size_t msum = 0, csum = 0;
  for (int32_t N : hd.poly_length) {
    msum += N;
    csum += (size_t)N * N;
  }
  hd.monotype.assign(desc->monotype, desc->monotype + msum);
  hd.connection.assign(desc->connection, desc->connection + csum);
  return create_impl(out, hd, opts);
}

int scmf_create_from_files(scmf_sys** out, const char* configuration_dat, const char* polymer_dat,
                           const scmf_opts* opts) {
  if (!configuration_dat || !polymer_dat) return fail(SCMF_E_ARG, "file path is null");
  HostDesc hd;
  std::string e = read_polymer(polymer_dat, &hd);
  if (!e.empty()) return fail(SCMF_E_IO, e);
  e = read_configuration(configuration_dat, &hd);
  if (!e.empty()) return fail(SCMF_E_IO, e);
  return create_impl(out, hd, opts);
}

void scmf_destroy(scmf_sys* s) {
  if (!s) return;
  (void)hipStreamSynchronize(s->stream);
  free_all(s);
  delete s;
}

int scmf_mc_sweep(scmf_sys* s, int32_t n_sweeps) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  if (n_sweeps < 0) return fail(SCMF_E_ARG, "n_sweeps must be non-negative");
  for (int32_t i = 0; i < n_sweeps; ++i) {
    Timed t;
    if (s->timing) {
      int rc = s->ev_begin(K_SWEEP, &t);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(k_sweep, dim3(grid_for(s->n_chains)), dim3(BLOCK), 0, s->stream, s->kp,
                       s->x, s->y, s->z, s->omega, (uint32_t)s->sweeps, s->d_accepts);
    HIP_TRY(hipGetLastError());
    if (s->timing) {
      int rc = s->ev_end(&t);
      if (rc) return rc;
    }
    s->sweeps += 1;
    s->moves += (uint64_t)s->n_beads;
  }
  return SCMF_OK;
}

int scmf_grid_density(scmf_sys* s) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  Timed t;
  if (s->timing) {
    int rc = s->ev_begin(K_GRID, &t);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_zero, dim3(grid_strided(s->count_words)), dim3(BLOCK), 0, s->stream,
                     s->counts, s->count_words);
  if (s->opts.density_bits == 32)
    hipLaunchKernelGGL(k_grid<32>, dim3(grid_strided(s->n_beads)), dim3(BLOCK), 0, s->stream,
                       s->kp, s->x, s->y, s->z, s->counts, s->d_overflow);
  else
    hipLaunchKernelGGL(k_grid<16>, dim3(grid_strided(s->n_beads)), dim3(BLOCK), 0, s->stream,
                       s->kp, s->x, s->y, s->z, s->counts, s->d_overflow);
  HIP_TRY(hipGetLastError());
  if (s->timing) return s->ev_end(&t);
  return SCMF_OK;
}

int scmf_update_omega(scmf_sys* s) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  Timed t;
  if (s->timing) {
    int rc = s->ev_begin(K_OMEGA, &t);
    if (rc) return rc;
  }
  if (s->opts.density_bits == 32)
    hipLaunchKernelGGL(k_omega<32>, dim3(grid_strided(s->n_cells)), dim3(BLOCK), 0, s->stream,
                       s->kp, s->counts, s->omega);
  else
    hipLaunchKernelGGL(k_omega<16>, dim3(grid_strided(s->n_cells)), dim3(BLOCK), 0, s->stream,
                       s->kp, s->counts, s->omega);
  HIP_TRY(hipGetLastError());
  if (s->timing) return s->ev_end(&t);
  return SCMF_OK;
}

int scmf_step(scmf_sys* s, int32_t step) {
  int rc = scmf_mc_sweep(s, 1);
  if (rc) return rc;
  if ((rc = scmf_grid_density(s))) return rc;
  if (s->opts.omega_every_step || (step % 10)) return scmf_update_omega(s);
  return SCMF_OK;
}

int scmf_copy_out(scmf_sys* s, int32_t what, void* host, size_t bytes) {
  if (!s || !host) return fail(SCMF_E_ARG, "null argument");
  const int64_t TC = (int64_t)s->T * s->n_cells;
  if (what == SCMF_POS) {
    if (bytes < sizeof(float) * 3 * (size_t)s->n_beads) return fail(SCMF_E_ARG, "buffer too small");
    int rc = ensure_staging(s, std::min<int64_t>(s->n_beads, STAGE_BEADS));
    if (rc) return rc;
    for (int64_t b0 = 0; b0 < s->n_beads; b0 += STAGE_BEADS) {
      const int64_t n = std::min<int64_t>(STAGE_BEADS, s->n_beads - b0);
      hipLaunchKernelGGL(k_planar_to_aos, dim3(grid_strided(n)), dim3(BLOCK), 0, s->stream,
                         s->x + b0, s->y + b0, s->z + b0, s->staging, n);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync((float*)host + 3 * b0, s->staging, sizeof(float) * 3 * n,
                             hipMemcpyDeviceToHost, s->stream));
      HIP_TRY(hipStreamSynchronize(s->stream));
    }
    return SCMF_OK;
  }
  if (what == SCMF_DENSITY) {
    if (bytes < sizeof(uint32_t) * (size_t)TC) return fail(SCMF_E_ARG, "buffer too small");
    if (s->opts.density_bits == 32) {
      HIP_TRY(hipMemcpyAsync(host, s->counts, sizeof(uint32_t) * TC, hipMemcpyDeviceToHost, s->stream));
    } else {
      uint32_t* wide = nullptr;
      HIP_TRY(hipMalloc(&wide, sizeof(uint32_t) * TC));
      hipLaunchKernelGGL(k_widen<16>, dim3(grid_strided(TC)), dim3(BLOCK), 0, s->stream, s->counts,
                         wide, TC);
      hipError_t e1 = hipMemcpyAsync(host, wide, sizeof(uint32_t) * TC, hipMemcpyDeviceToHost, s->stream);
      hipError_t e2 = hipStreamSynchronize(s->stream);
      (void)hipFree(wide);
      HIP_TRY(e1);
      HIP_TRY(e2);
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SCMF_OK;
  }
  if (what == SCMF_OMEGA) {
    if (bytes < sizeof(float) * (size_t)TC) return fail(SCMF_E_ARG, "buffer too small");
    HIP_TRY(hipMemcpyAsync(host, s->omega, sizeof(float) * TC, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SCMF_OK;
  }
  return fail(SCMF_E_ARG, "unknown copy_out selector");
}

int scmf_copy_in_pos(scmf_sys* s, const float* xyz, size_t n_beads) {
  if (!s || !xyz) return fail(SCMF_E_ARG, "null argument");
  if ((int64_t)n_beads != s->n_beads) return fail(SCMF_E_ARG, "bead count mismatch");
  int rc = ensure_staging(s, std::min<int64_t>(s->n_beads, STAGE_BEADS));
  if (rc) return rc;
  for (int64_t b0 = 0; b0 < s->n_beads; b0 += STAGE_BEADS) {
    const int64_t n = std::min<int64_t>(STAGE_BEADS, s->n_beads - b0);
    HIP_TRY(hipMemcpyAsync(s->staging, xyz + 3 * b0, sizeof(float) * 3 * n, hipMemcpyHostToDevice,
                           s->stream));
    hipLaunchKernelGGL(k_aos_to_planar, dim3(grid_strided(n)), dim3(BLOCK), 0, s->stream,
                       s->staging, s->x + b0, s->y + b0, s->z + b0, n);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
  }
  return SCMF_OK;
}

int scmf_copy_in_omega(scmf_sys* s, const float* omega, size_t n_values) {
  if (!s || !omega) return fail(SCMF_E_ARG, "null argument");
  const int64_t TC = (int64_t)s->T * s->n_cells;
  if ((int64_t)n_values != TC) return fail(SCMF_E_ARG, "omega size mismatch");
  HIP_TRY(hipMemcpyAsync(s->omega, omega, sizeof(float) * TC, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return SCMF_OK;
}

int scmf_copy_in_density(scmf_sys* s, const uint32_t* counts, size_t n_values) {
  if (!s || !counts) return fail(SCMF_E_ARG, "null argument");
  const int64_t TC = (int64_t)s->T * s->n_cells;
  if ((int64_t)n_values != TC) return fail(SCMF_E_ARG, "density size mismatch");
  if (s->opts.density_bits == 32) {
    HIP_TRY(hipMemcpyAsync(s->counts, counts, sizeof(uint32_t) * TC, hipMemcpyHostToDevice, s->stream));
  } else {
    std::vector<uint32_t> packed(s->count_words, 0u);
    for (int64_t i = 0; i < TC; ++i) {
      if (counts[i] > 0xFFFFu) return fail(SCMF_E_OVERFLOW, "count above 65535 for uint16 fields");
      packed[i >> 1] |= counts[i] << ((i & 1) * 16);
    }
    HIP_TRY(hipMemcpyAsync(s->counts, packed.data(), sizeof(uint32_t) * s->count_words,
                           hipMemcpyHostToDevice, s->stream));
  }
  HIP_TRY(hipStreamSynchronize(s->stream));
  return SCMF_OK;
}

int scmf_set_rng(scmf_sys* s, uint64_t seed, uint64_t next_sweep) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  if (next_sweep >= 0xFFFFFFFFull) return fail(SCMF_E_ARG, "sweep counter exhausted");
  s->opts.seed = seed;
  s->kp.seed = seed;
  s->sweeps = next_sweep;
  return SCMF_OK;
}

int scmf_stats(scmf_sys* s, scmf_stats_t* out) {
  if (!s || !out) return fail(SCMF_E_ARG, "null argument");
  const unsigned nb = grid_for(s->n_chains);
  hipLaunchKernelGGL(k_re2, dim3(nb), dim3(BLOCK), 0, s->stream, s->kp, s->x, s->y, s->z,
                     s->d_partial);
  HIP_TRY(hipGetLastError());
  std::vector<double> part(nb);
  unsigned long long acc = 0;
  uint32_t ovf = 0;
  HIP_TRY(hipMemcpyAsync(part.data(), s->d_partial, sizeof(double) * nb, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipMemcpyAsync(&acc, s->d_accepts, sizeof acc, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipMemcpyAsync(&ovf, s->d_overflow, sizeof ovf, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  double sum = 0;
  for (double v : part) sum += v;
  out->moves = s->moves;
  out->accepts = acc;
  out->sweeps = s->sweeps;
  out->re2 = sum / (double)s->n_chains;
  out->overflow = ovf ? 1 : 0;
  return SCMF_OK;
}

int scmf_reset_stats(scmf_sys* s) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  HIP_TRY(hipMemsetAsync(s->d_accepts, 0, sizeof(unsigned long long), s->stream));
  HIP_TRY(hipMemsetAsync(s->d_overflow, 0, sizeof(uint32_t), s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  s->moves = 0;
  return SCMF_OK;
}

int scmf_enable_timing(scmf_sys* s, int32_t on) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  s->timing = on != 0;
  return SCMF_OK;
}

int scmf_timing(scmf_sys* s, scmf_timing_t* out) {
  if (!s || !out) return fail(SCMF_E_ARG, "null argument");
  int rc = s->drain_timing();
  if (rc) return rc;
  out->sweep_n = s->t_n[K_SWEEP];
  out->grid_n = s->t_n[K_GRID];
  out->omega_n = s->t_n[K_OMEGA];
  out->sweep_ms = s->t_n[K_SWEEP] ? s->t_sum[K_SWEEP] / s->t_n[K_SWEEP] : 0;
  out->grid_ms = s->t_n[K_GRID] ? s->t_sum[K_GRID] / s->t_n[K_GRID] : 0;
  out->omega_ms = s->t_n[K_OMEGA] ? s->t_sum[K_OMEGA] / s->t_n[K_OMEGA] : 0;
  return SCMF_OK;
}

int scmf_reset_timing(scmf_sys* s) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  int rc = s->drain_timing();
  if (rc) return rc;
  for (int k = 0; k < K_KINDS; ++k) {
    s->t_sum[k] = 0;
    s->t_n[k] = 0;
  }
  return SCMF_OK;
}

int scmf_sync(scmf_sys* s) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  HIP_TRY(hipStreamSynchronize(s->stream));
  return SCMF_OK;
}

int scmf_sizes(scmf_sys* s, int64_t* n_beads, int64_t* n_chains, int64_t* n_cells, int32_t* n_types) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  if (n_beads) *n_beads = s->n_beads;
  if (n_chains) *n_chains = s->n_chains;
  if (n_cells) *n_cells = s->n_cells;
  if (n_types) *n_types = s->T;
  return SCMF_OK;
}

int scmf_write_density(scmf_sys* s, const char* dir, int32_t step) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  const int64_t TC = (int64_t)s->T * s->n_cells;
  std::vector<uint32_t> counts(TC);
  std::vector<float> om(TC);
  int rc = scmf_copy_out(s, SCMF_DENSITY, counts.data(), sizeof(uint32_t) * TC);
  if (rc) return rc;
  if ((rc = scmf_copy_out(s, SCMF_OMEGA, om.data(), sizeof(float) * TC))) return rc;
  std::string path = density_file_name(step);
  if (dir && *dir) path = std::string(dir) + "/" + path;
  std::string e = write_density_file(path.c_str(), s->desc.n, s->desc.L, s->T, s->n_cells,
                                     counts.data(), om.data(), s->kp.fst[0]);
  if (!e.empty()) return fail(SCMF_E_IO, e);
  return SCMF_OK;
}

int scmf_write_coord(scmf_sys* s, const char* path) {
  if (!s || !path) return fail(SCMF_E_ARG, "null argument");
  std::vector<float> xyz(3 * (size_t)s->n_beads);
  int rc = scmf_copy_out(s, SCMF_POS, xyz.data(), sizeof(float) * xyz.size());
  if (rc) return rc;
  std::string e = write_coord_file(path, xyz.data(), s->n_beads);
  if (!e.empty()) return fail(SCMF_E_IO, e);
  return SCMF_OK;
}

namespace {
void fill_summary(const scmf_desc* d, scmf_summary* out) {
  memset(out, 0, sizeof *out);
  out->n_mono_types = d->n_mono_types;
  out->n_poly_types = d->n_poly_types;
  out->reference_Nbeads = d->reference_Nbeads;
  for (int i = 0; i < 3; ++i) {
    out->n[i] = d->n[i];
    out->L[i] = d->L[i];
  }
  out->dtNxi = d->dtNxi;
  int64_t chains = 0, beads = 0;
  for (int p = 0; p < d->n_poly_types; ++p) {
    chains += d->n_polymers[p];
    beads += d->n_polymers[p] * (int64_t)d->poly_length[p];
  }
  out->n_chains = chains;
  out->n_beads = beads;
  out->n_cells = (int64_t)d->n[0] * d->n[1] * d->n[2];
  const int tt = d->n_mono_types * d->n_mono_types;
  for (int i = 0; i < tt && i < 64; ++i) out->xn[i] = d->xn[i];
  const uint64_t TC = (uint64_t)d->n_mono_types * out->n_cells;
  const uint64_t fixed = 12ull * beads + 4ull * TC + 8ull * ((chains + BLOCK - 1) / BLOCK) + 4096;
  out->device_bytes_16 = fixed + 4ull * ((TC + 1) / 2);
  out->device_bytes_32 = fixed + 4ull * TC;
}
}  // namespace

int scmf_inspect_desc(const scmf_desc* desc, scmf_summary* out) {
  if (!desc || !out) return fail(SCMF_E_ARG, "null argument");
  std::string v = validate_desc(desc);
  if (!v.empty()) return fail(SCMF_E_ARG, v);
  fill_summary(desc, out);
  return SCMF_OK;
}

int scmf_inspect_files(const char* configuration_dat, const char* polymer_dat, scmf_summary* out) {
  if (!configuration_dat || !polymer_dat || !out) return fail(SCMF_E_ARG, "null argument");
  HostDesc hd;
  std::string e = read_polymer(polymer_dat, &hd);
  if (!e.empty()) return fail(SCMF_E_IO, e);
  e = read_configuration(configuration_dat, &hd);
  if (!e.empty()) return fail(SCMF_E_IO, e);
  scmf_desc d = hd.view();
  return scmf_inspect_desc(&d, out);
}

int scmf_write_density_host(const char* path, const int32_t n[3], const double L[3], int32_t T,
                            const uint32_t* counts, const float* omega, double field_scaling) {
  if (!path || !n || !L || !counts || !omega || T <= 0) return fail(SCMF_E_ARG, "null argument");
  const int64_t C = (int64_t)n[0] * n[1] * n[2];
  std::string e = write_density_file(path, n, L, T, C, counts, omega, field_scaling);
  if (!e.empty()) return fail(SCMF_E_IO, e);
  return SCMF_OK;
}

int scmf_density_to_vtk(const char* density_path, const char* vtk_path) {
  if (!density_path || !vtk_path) return fail(SCMF_E_ARG, "null path");
  std::string e = density_to_vtk(density_path, vtk_path);
  if (!e.empty()) return fail(SCMF_E_IO, e);
  return SCMF_OK;
}

}  // extern "C"
