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

// Global-address-space pointers: loads through them are vector-memory
// instructions only (a pointer read from memory would otherwise be flat and
// also wait on the LDS counter).
#define SCMF_GLOBAL __attribute__((address_space(1)))
// Constant address space: the kernels read the system parameters through it,
// so the loads are invariant (hoisted across barriers) and scalar when uniform.
#define SCMF_CONST __attribute__((address_space(4)))

namespace scmf {

template <class T>
__device__ __forceinline__ SCMF_GLOBAL T* gptr(T* p) {
  return (SCMF_GLOBAL T*)p;
}

constexpr int MAX_T = 8;   // monomer types
constexpr int MAX_P = 16;  // polymer architectures

// Everything a kernel needs about one system, passed by value.
struct KParams {
  int T, P;
  int nx, ny, nz;
  uint32_t n_cells;
  double Lx, Ly, Lz, dx, dy, dz;
  Axis ax[3];                    // binning constants per axis
  int pow2;                      // every axis has power-of-two cells
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
  int type_bits;                 // 1: at most two monomer types and 64 beads a chain, so
  uint64_t tbits[MAX_P];         //    bead b of architecture p has type (tbits[p] >> b) & 1
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
  unsigned long long* accepts;    // ACC_SLOTS partial counts (summed on the host)
  uint32_t* overflow;
  uint32_t* ctr;                 // [0] stream sweep word of the next sweep
  uint32_t* cells;               // packed cell of every stored bead, written by the sweep
                                 // (x | y << 10 | z << 20), or null
  uint8_t* sched;                // per slot: the sweep's move schedule made by the schedule kernels
  int sched_stride;              // bytes per slot record
  uint32_t stop;                 // batches: the member does no work once its sweep word
                                 // reaches this (NO_STOP: never); ctr[1] says whether the
                                 // current step's sweep ran
};

constexpr uint32_t NO_STOP = 0xFFFFFFFFu;

// Accepted moves are counted in this many separate words, chosen by
// workgroup, so the workgroups of a sweep do not queue on one address.
constexpr int ACC_SLOTS = 256;

// Which system and which of its blocks a workgroup works on.  Kernels that
// serve a batch of systems get a table with one entry per workgroup; a single
// system passes none.
struct BlockRef {
  uint32_t sys, local, nblk, pad;
};

using KPC = const SCMF_CONST KParams;

// The parameters of system `sys` seen through the constant address space.
__device__ __forceinline__ KPC& param(const KParams* Ps, uint32_t sys) {
  return *(KPC*)(Ps + sys);
}

// A batch member whose sweep word has reached its stop sweeps no more.
template <class KP>
__device__ __forceinline__ bool sweep_stopped(const KP& kp) {
  return ((const uint32_t*)kp.ctr)[0] >= kp.stop;
}

// The gridding and omega update of a step run only if the step's sweep ran
// (k_advance sets ctr[1]); a member without a stop always runs them, also
// outside the step sequence.
template <class KP>
__device__ __forceinline__ bool step_skipped(const KP& kp) {
  return kp.stop != NO_STOP && ((const uint32_t*)kp.ctr)[1] == 0u;
}

// Inside scmf_step two small launches ride on the step's other kernels: the
// sweep clears the counts the step's gridding then fills (nothing reads them
// in between), and the gridding advances the stream sweep word (nothing reads
// it before the next sweep).  clear_counts runs in every workgroup of a sweep
// launch, advance_sweep_word in one lane of a gridding launch.
template <class KP>
__device__ __forceinline__ void clear_counts(const KP& kp) {
  SCMF_GLOBAL uint32_t* __restrict__ cw = gptr((uint32_t*)kp.counts);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < kp.count_words; q += stride)
    cw[q] = 0u;
}
template <class KP>
__device__ __forceinline__ void advance_sweep_word(const KP& kp) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    uint32_t* c = (uint32_t*)kp.ctr;
    const bool ran = c[0] < kp.stop;
    if (ran) c[0] = c[0] + 1u;
    c[1] = ran ? 1u : 0u;
  }
}

__device__ __forceinline__ BlockRef block_ref(const BlockRef* bt) {
  if (bt) return bt[blockIdx.x];
  BlockRef r;
  r.sys = 0;
  r.local = blockIdx.x;
  r.nblk = gridDim.x;
  r.pad = 0;
  return r;
}



// Binning constants of one axis, copied out of the parameters (whatever
// address space they are read from).
template <class KP>
__device__ __forceinline__ Axis axis(const KP& kp, int a) {
  Axis r;
  r.L = kp.ax[a].L;
  r.d = kp.ax[a].d;
  r.invL = kp.ax[a].invL;
  r.invd = kp.ax[a].invd;
  r.guard2 = kp.ax[a].guard2;
  r.lhi = kp.ax[a].lhi;
  r.gin = kp.ax[a].gin;
  r.gin1 = kp.ax[a].gin1;
  r.invd_f = kp.ax[a].invd_f;
  r.lhi_f = kp.ax[a].lhi_f;
  r.n = kp.ax[a].n;
  r.pow2 = kp.ax[a].pow2;
  r.wrap2 = kp.ax[a].wrap2;
  r.rneg_f = kp.ax[a].rneg_f;
  r.gw1 = kp.ax[a].gw1;
  return r;
}

// Linear index x + nx (y + ny z) of a cell (device_function.cuh:19).
template <class KP>
__device__ __forceinline__ uint32_t cell_index(const KP& kp, int xi, int yi, int zi) {
  return (uint32_t)(xi + yi * kp.nx + zi * kp.nx * kp.ny);
}

// NGP cell (within one type's field) of a position, coordinates given in
// fp32 (a stored position) or fp64 (a sum the reference forms in fp64).
template <class KP, class R>
__device__ __forceinline__ uint32_t cell_of(const KP& kp, R rx, R ry, R rz) {
  const Axis ax = axis(kp, 0), ay = axis(kp, 1), az = axis(kp, 2);
  int xi, yi, zi;
  const bool ok = kp.pow2 && (fold_pow2(rx, ax, &xi) & fold_pow2(ry, ay, &yi) &
                              fold_pow2(rz, az, &zi));
  if (!ok) {
    xi = fold((double)rx, ax);
    yi = fold((double)ry, ay);
    zi = fold((double)rz, az);
  }
  return cell_index(kp, xi, yi, zi);
}

template <class KP, class R>
__device__ __forceinline__ uint32_t bin_cell(const KP& kp, R rx, R ry, R rz, uint32_t type) {
  return type * kp.n_cells + cell_of(kp, rx, ry, rz);
}

// polymer_number_to_index (device_function.cu:72-93): architecture of a chain.
template <class KP>
__device__ __forceinline__ int arch_of_chain(const KP& kp, int64_t chain) {
  int p = 0;
  while (p < kp.P - 1 && chain >= kp.chain_base[p + 1]) ++p;
  return p;
}

template <class KP>
__device__ __forceinline__ int arch_of_bead(const KP& kp, int64_t bead) {
  int p = 0;
  while (p < kp.P - 1 && bead >= kp.bead_base[p + 1]) ++p;
  return p;
}

}  // namespace scmf
