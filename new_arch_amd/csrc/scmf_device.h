// scmf_device.h — device-side system description and the per-bead arithmetic
// of the SCMF Monte Carlo path, written for gfx950 (wave64).
//
// The arithmetic mirrors the reference operation for operation (cited per
// function) so results are bit-identical to the CPU oracle; the data layout is
// this engine's own: planar fp32 coordinates x[], y[], z[] with every chain's
// beads contiguous, counts as [type][z][y][x].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "scmf_bin.h"
#include "scmf_prims.h"

#pragma clang fp contract(off)

namespace scmf {

constexpr int MAX_T = 8;   // monomer types
constexpr int MAX_P = 16;  // polymer architectures

// Everything a kernel needs about one system, passed by value.
struct KParams {
  int T, P;
  int nx, ny, nz;
  uint32_t n_cells;
  double Lx, Ly, Lz, dx, dy, dz;
  Axis ax[3];                    // binning constants per axis
  double harm;                   // harmonic_normb
  double inv_nref;               // inverse_refbeads
  double fst[MAX_T];             // field_scaling_type
  double xn[MAX_T * MAX_T];      // [t + T*s]
  float A[MAX_T], R[MAX_T];      // smart-MC step, as the move reads them (fp32)
  int64_t chain_base[MAX_P + 1]; // polymer_basis_gpu
  int64_t bead_base[MAX_P + 1];  // monomer_poly_basis_gpu
  int len[MAX_P];
  int linear[MAX_P];             // 1: bonds are exactly i <-> i+1 (neighbours computed, not looked up)
  double inv_len[MAX_P];         // 1/len, for the bead -> (chain, monomer) split
  int mono_off[MAX_P + 1];
  const int* monotype;           // [sum N_p]
  const int* nbr_start;          // [sum N_p + 1], CSR over every bead of every type
  const int* nbr;                // neighbour bead indices within the chain
  const uint32_t* slot_chain;    // chain id stored in each slot (slots ordered spatially)
  const uint32_t* chain_slot;    // slot of each chain id
  const uint8_t* btype;          // monomer type of every stored bead
  int64_t n_chains, n_beads;
  uint64_t seed;
  uint32_t system_id;
  int move_type;
  float trial_delta;
  // device state of the system
  float* x;
  float* y;
  float* z;
  float* omega;
  uint32_t* counts;              // uint16 pairs packed in words, or uint32
  int64_t count_words;
  int density_bits;
  int omega_every_step;
  unsigned long long* accepts;
  uint32_t* overflow;
  uint32_t* ctr;                 // [0] stream sweep word of the next sweep, [1] finished sweep blocks
};

// Which system and which of its blocks a workgroup works on.  Kernels that
// serve a batch of systems get a table with one entry per workgroup; a single
// system passes none.
struct BlockRef {
  uint32_t sys, local, nblk, pad;
};

__device__ __forceinline__ BlockRef block_ref(const BlockRef* bt) {
  if (bt) return bt[blockIdx.x];
  BlockRef r;
  r.sys = 0;
  r.local = blockIdx.x;
  r.nblk = gridDim.x;
  r.pad = 0;
  return r;
}

// End of a sweep launch for one system: the last of its workgroups to finish
// advances the system's sweep word (every workgroup has read it by then).
__device__ __forceinline__ void finish_sweep_block(const KParams& kp, const BlockRef& br) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&kp.ctr[1], 1u) == br.nblk - 1) {
      atomicExch(&kp.ctr[1], 0u);
      atomicAdd(&kp.ctr[0], 1u);
    }
  }
}

__device__ __forceinline__ uint32_t bin_cell(const KParams& kp, double rx, double ry, double rz,
                                             uint32_t type) {
  const int xi = fold_fast(rx, kp.ax[0]);
  const int yi = fold_fast(ry, kp.ax[1]);
  const int zi = fold_fast(rz, kp.ax[2]);
  return type * kp.n_cells + (uint32_t)(xi + yi * kp.nx + zi * kp.nx * kp.ny);
}

// polymer_number_to_index (device_function.cu:72-93): architecture of a chain.
__device__ __forceinline__ int arch_of_chain(const KParams& kp, int64_t chain) {
  int p = 0;
  while (p < kp.P - 1 && chain >= kp.chain_base[p + 1]) ++p;
  return p;
}

__device__ __forceinline__ int arch_of_bead(const KParams& kp, int64_t bead) {
  int p = 0;
  while (p < kp.P - 1 && bead >= kp.bead_base[p + 1]) ++p;
  return p;
}

}  // namespace scmf
