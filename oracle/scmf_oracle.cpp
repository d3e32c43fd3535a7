// scmf_oracle.cpp — CPU restatement of the reference's SCMF Monte Carlo hot path.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP engine is
// compared against, and the CPU baseline timed beside it.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the engine
// never calls it.
//
// Each function restates one reference function (file:line cited) in plain
// C++, keeping the reference's precision mix: fp32 positions, forces and
// displacements; fp64 binning, energies and acceptance.  Compile with FP
// contraction off (see oracle/Makefile) so every operation rounds once.
// The reference's random numbers (cuRAND) and exp are replaced by the injected
// streams of new_arch_amd/csrc/scmf_prims.h; those primitives are pinned
// separately (Philox known-answer vectors, libm error bounds).
//
// Parity pins (tests/test_oracle_pins.py, tests/test_prims.py):
//   binning      exact rational arithmetic, hand-derived edge cases, periodicity,
//                invariants (sum of counts, per-type sums, order independence);
//   omega        the reference's own density_*.dat files: their omega columns are
//                reproduced from the integer counts they decode to; closed forms;
//   derived      harm = 3 (Nref - 1) / 2, A = dtNxi / Nref, R = sqrt(2 A) with
//                N != Nref;
//   move         single moves of a 3-bead chain with N != Nref computed by hand
//                from the bond energy, the force-biased Gaussian proposal and the
//                Metropolis-Hastings ratio (smart and trial moves, the uniform just
//                above and just below exp(-dE), the drift and the noise separately,
//                the fp64 new cell across a cell edge, the bead pick); relaxation
//                to the ideal chain with harm from Nref (N = Nref and N != Nref);
//                detailed balance of dimers (chi-square of the bond length);
//                Boltzmann occupancy of a two-level field;
//   init walk    bond variance 1/(3 (N - 1)) from the chain length with N != Nref,
//                <Re^2> = 1;
//   streams      Philox known-answer vectors, special functions against libm.
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../include/scmf.h"
#include "../new_arch_amd/csrc/scmf_prims.h"

#ifdef _OPENMP
#include <omp.h>
#endif

using namespace scmf;

namespace {

// Derived system constants: init_config.cu:480-484 (harmonic_normb),
// :565-567 (dx), :662 (inverse_refbeads), :670-672 (xn layout),
// :784-785 (field_scaling_type), :804-809 (A, R).
struct Sys {
  int T;
  int nx, ny, nz;
  double Lx, Ly, Lz, dx, dy, dz;
  int64_t n_cells;
  std::vector<double> xn;  // [t + T*s]
  double inverse_refbeads;
  double harmonic_normb;
  std::vector<double> field_scaling_type;
  std::vector<double> A, R;
  int P;
  std::vector<int> len;
  std::vector<int64_t> chain_base, bead_base;  // polymer_basis_gpu, monomer_poly_basis_gpu
  std::vector<int> mono_off;                    // offset of type p in monotype
  std::vector<int> monotype;
  std::vector<int> nbr_start;                   // CSR over all beads of all types
  std::vector<int> nbr;
  int64_t n_chains, n_beads;
};

bool build(const scmf_desc* d, Sys& s) {
  if (!d || d->n_mono_types <= 0 || d->n_poly_types <= 0) return false;
  s.T = d->n_mono_types;
  s.nx = d->n[0];
  s.ny = d->n[1];
  s.nz = d->n[2];
  s.Lx = d->L[0];
  s.Ly = d->L[1];
  s.Lz = d->L[2];
  s.dx = s.Lx / s.nx;
  s.dy = s.Ly / s.ny;
  s.dz = s.Lz / s.nz;
  s.n_cells = (int64_t)s.nx * s.ny * s.nz;
  s.xn.assign(d->xn, d->xn + s.T * s.T);
  const int nref = d->reference_Nbeads;
  s.inverse_refbeads = 1 / (double)nref;
  const double cste = 1.0 / sqrt(3.0 * ((double)nref - 1.0));
  s.harmonic_normb = 1.0 / (2.0 * cste * cste);
  s.P = d->n_poly_types;
  s.len.assign(d->poly_length, d->poly_length + s.P);
  s.chain_base.assign(s.P + 1, 0);
  s.bead_base.assign(s.P + 1, 0);
  s.mono_off.assign(s.P + 1, 0);
  for (int p = 0; p < s.P; ++p) {
    s.chain_base[p + 1] = s.chain_base[p] + d->n_polymers[p];
    s.bead_base[p + 1] = s.bead_base[p] + d->n_polymers[p] * (int64_t)s.len[p];
    s.mono_off[p + 1] = s.mono_off[p] + s.len[p];
  }
  s.n_chains = s.chain_base[s.P];
  s.n_beads = s.bead_base[s.P];
  s.monotype.assign(d->monotype, d->monotype + s.mono_off[s.P]);
  // Neighbour lists: bead i's neighbours are the j with connection[i][j] set,
  // ascending j (init_config.cu:312-336).
  s.nbr_start.assign(s.mono_off[s.P] + 1, 0);
  s.nbr.clear();
  int64_t conn_off = 0;
  for (int p = 0; p < s.P; ++p) {
    const int N = s.len[p];
    for (int i = 0; i < N; ++i) {
      s.nbr_start[s.mono_off[p] + i] = (int)s.nbr.size();
      for (int j = 0; j < N; ++j)
        if (d->connection[conn_off + (int64_t)i * N + j] == 1) s.nbr.push_back(j);
    }
    conn_off += (int64_t)N * N;
  }
  s.nbr_start[s.mono_off[s.P]] = (int)s.nbr.size();
  s.field_scaling_type.assign(s.T, (double)s.n_cells / (double)s.n_beads);
  s.A.assign(s.T, d->dtNxi / nref);
  s.R.assign(s.T, 0.0);
  for (int t = 0; t < s.T; ++t) s.R[t] = sqrt(s.A[t] * 2);
  return true;
}

// coord_to_index_unified_high, device_function.cu:313-360.  Coordinates the
// reference cannot bin (non-finite, or beyond 2^31 box lengths where its int
// conversions overflow) are put in cell 0, as the engine does.
inline int fold(double r, double L, double d, int n) {
  int i = 0;
  if (r >= 0) {
    const double q = r / L;
    if (!(q < 2147483647.0)) return 0;
    int k = (int)q;
    i = (int)((r - L * k) / d);
    i = i % n;
  } else if (r < 0) {
    const double q = -r / L;
    if (!(q < 2147483646.0)) return 0;
    int k = (int)q + 1;
    i = (int)((k * L + r) / d);
    i = i % n;
  }
  if ((unsigned)i >= (unsigned)n) i = 0;
  return i;
}

inline unsigned bin(const Sys& s, double rx, double ry, double rz, unsigned type) {
  const int xi = fold(rx, s.Lx, s.dx, s.nx);
  const int yi = fold(ry, s.Ly, s.dy, s.ny);
  const int zi = fold(rz, s.Lz, s.dz, s.nz);
  return (unsigned)(type * (unsigned)s.n_cells + xi + yi * s.nx + zi * s.nx * s.ny);
}

// polymer_number_to_index, device_function.cu:72-93 (linear search).
inline int poly_type_of(const Sys& s, int64_t chain) {
  for (int p = 0; p < s.P; ++p)
    if (chain < s.chain_base[p + 1]) return p;
  return s.P - 1;
}

// polymer_number_to_pos(id, 0), device_function.cu:104-133.
inline int64_t first_bead(const Sys& s, int p, int64_t chain) {
  return s.bead_base[p] + (chain - s.chain_base[p]) * s.len[p];
}

// add_bond_forces_arr, device_function.cu:589-634: the neighbour difference
// is formed in fp32, scaled and accumulated in fp64, added to the fp32 force.
inline void bond_forces(const Sys& s, int p, int mono, int64_t poly_index, const float* pos,
                        float x, float y, float z, float* fx, float* fy, float* fz) {
  const double harm = s.harmonic_normb;
  double v1x = 0.0, v1y = 0.0, v1z = 0.0;
  const int b = s.mono_off[p] + mono;
  for (int k = s.nbr_start[b]; k < s.nbr_start[b + 1]; ++k) {
    const int64_t ni = (poly_index + s.nbr[k]) * 3;
    v1x += (double)(float)(pos[ni] - x) * 2.0 * harm;
    v1y += (double)(float)(pos[ni + 1] - y) * 2.0 * harm;
    v1z += (double)(float)(pos[ni + 2] - z) * 2.0 * harm;
  }
  *fx = (float)((double)*fx + v1x);
  *fy = (float)((double)*fy + v1y);
  *fz = (float)((double)*fz + v1z);
}

// calc_delta_bonded_energy_arr, device_function.cu:509-561.
inline double bonded_delta(const Sys& s, int p, int mono, int64_t poly_index, const float* pos,
                           double dx, double dy, double dz) {
  double delta_energy = 0;
  const int b = s.mono_off[p] + mono;
  const int64_t index = poly_index + mono;
  for (int k = s.nbr_start[b]; k < s.nbr_start[b + 1]; ++k) {
    const int64_t neigh_index = poly_index + s.nbr[k];
    const double old_rx = (double)(float)(pos[index * 3] - pos[neigh_index * 3]);
    const double new_rx = old_rx + dx;
    const double old_ry = (double)(float)(pos[index * 3 + 1] - pos[neigh_index * 3 + 1]);
    const double new_ry = old_ry + dy;
    const double old_rz = (double)(float)(pos[index * 3 + 2] - pos[neigh_index * 3 + 2]);
    const double new_rz = old_rz + dz;
    const double old_r2 = old_rx * old_rx + old_ry * old_ry + old_rz * old_rz;
    const double new_r2 = new_rx * new_rx + new_ry * new_ry + new_rz * new_rz;
    delta_energy += s.harmonic_normb * (new_r2 - old_r2);
  }
  return delta_energy;
}

// mc_polymer_move_arr, device_function.cu:718-799, for one chain, with
// trial_move_smc_arr (:637-717), calc_delta_energy_arr (:562-588) and
// som_accept (:444-454) inlined.  Returns the number of accepted moves.
// The random numbers of move i of a chain: the Philox streams the engine
// uses, or arrays injected by the caller (any other generator).
// Move i reads its pick from the SLOT_PICK block of index i / 2 (words 0, 1
// for even i, 2, 3 for odd i) and its displacement and accept uniform from
// the SLOT_DISP block of index i (scmf_prims.h).
struct PhiloxDraws {
  const scmf_opts& o;
  uint32_t sweep;
  int64_t chain;
  u32x4 pick, disp;
  uint32_t move = 0;
  void at(int i) {
    move = (uint32_t)i;
    pick = stream_block(o.seed, o.system_id, (uint32_t)chain, sweep, (uint32_t)i >> 1, SLOT_PICK);
    disp = stream_block(o.seed, o.system_id, (uint32_t)chain, sweep, (uint32_t)i, SLOT_DISP);
  }
  double pick_u() const { return pick_u01(pick, move); }
  void normals(float* rx, float* ry, float* rz) const { normal3f(disp, rx, ry, rz); }
  void trial(float* ux, float* uy, float* uz) const {
    *ux = u01_24f_closed0(disp.v[0]);
    *uy = u01_24f_closed0(disp.v[1]);
    *uz = u01_24f_closed0(disp.v[2]);
  }
  float accept_u() const { return accept_u01(disp); }
};

// Injected draws, indexed by move: move i of the chain whose first bead is
// b0 reads pick[b0+i] in [0,1] (1 picks the last bead), gauss[3(b0+i)..] (three normals, or three
// uniforms in [0,1) for trial moves) and acc[b0+i] in (0,1].
struct InjectedDraws {
  const double* pick_;
  const float* gauss_;
  const float* acc_;
  int64_t m = 0;
  int64_t b0 = 0;
  void at(int i) { m = b0 + i; }
  double pick_u() const { return pick_[m]; }
  void normals(float* rx, float* ry, float* rz) const {
    *rx = gauss_[3 * m];
    *ry = gauss_[3 * m + 1];
    *rz = gauss_[3 * m + 2];
  }
  void trial(float* ux, float* uy, float* uz) const { normals(ux, uy, uz); }
  float accept_u() const { return acc_[m]; }
};

template <class Draws>
uint64_t chain_sweep(const Sys& s, const scmf_opts& o, float* pos, const float* omega,
                     int64_t chain, Draws& rnd) {
  const int p = poly_type_of(s, chain);
  const int64_t poly_index = first_bead(s, p, chain);
  const int length = s.len[p];
  uint64_t acc = 0;
  for (int i = 0; i < length; i++) {
    rnd.at(i);
    int ibead = (int)(rnd.pick_u() * length);
    if (ibead == length) ibead = length - 1;
    const unsigned iwtype = (unsigned)s.monotype[s.mono_off[p] + ibead];
    const int64_t index = poly_index + ibead;
    float dx, dy, dz;
    double smc_deltaE = 0.0;
    if (o.move_type == SCMF_MOVE_SMART) {
      float x = pos[index * 3], y = pos[index * 3 + 1], z = pos[index * 3 + 2];
      const float A = (float)s.A[iwtype];
      const float R = (float)s.R[iwtype];
      float fx = 0.0f, fy = 0.0f, fz = 0.0f;
      bond_forces(s, p, ibead, poly_index, pos, x, y, z, &fx, &fy, &fz);
      float rx, ry, rz;
      rnd.normals(&rx, &ry, &rz);
      dx = A * fx + rx * R;
      dy = A * fy + ry * R;
      dz = A * fz + rz * R;
      x += dx;
      y += dy;
      z += dz;
      float nfx = 0.0f, nfy = 0.0f, nfz = 0.0f;
      bond_forces(s, p, ibead, poly_index, pos, x, y, z, &nfx, &nfy, &nfz);
      smc_deltaE = 0.0;
      smc_deltaE += 0.5 * ((nfx + fx) * dx + (nfy + fy) * dy + (nfz + fz) * dz);
      smc_deltaE += 0.25 * A * ((nfx * nfx) + (nfy * nfy) + (nfz * nfz) - (fx * fx) - (fy * fy) -
                                (fz * fz));
    } else {
      const float delta = o.trial_delta;
      float ux, uy, uz;
      rnd.trial(&ux, &uy, &uz);
      dx = delta * (2.0f * ux - 1.0f);
      dy = delta * (2.0f * uy - 1.0f);
      dz = delta * (2.0f * uz - 1.0f);
    }
    const float newx = pos[index * 3] + dx;
    const float newy = pos[index * 3 + 1] + dy;
    const float newz = pos[index * 3 + 2] + dz;
    // calc_delta_energy_arr: old cell from the fp32 position, new cell from
    // the fp64 sum old + d (:573-577).
    const double xold = pos[index * 3], yold = pos[index * 3 + 1], zold = pos[index * 3 + 2];
    const unsigned cell_old = bin(s, xold, yold, zold, iwtype);
    const double energy_old = omega[cell_old];
    const unsigned cell_new = bin(s, xold + (double)dx, yold + (double)dy, zold + (double)dz,
                                  iwtype);
    const double energy_new = omega[cell_new];
    const double bonded = bonded_delta(s, p, ibead, poly_index, pos, dx, dy, dz);
    double energy = energy_new - energy_old;
    energy += bonded;
    double delta_energy = energy;
    delta_energy += smc_deltaE;
    // som_accept: the uniform matters only when p <= 1 (:449).
    const double p_acc = exp_cw(-1.0 * delta_energy);
    const bool accept = (p_acc > 1) || (p_acc > (double)rnd.accept_u());
    if (accept) {
      pos[index * 3] = newx;
      pos[index * 3 + 1] = newy;
      pos[index * 3 + 2] = newz;
      ++acc;
    }
  }
  return acc;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- derived constants
typedef struct orc_derived {
  double dx, dy, dz;
  double inverse_refbeads;
  double harmonic_normb;
  double field_scaling;
  double A, R;
  int64_t n_cells, n_chains, n_beads;
} orc_derived;

int orc_derive(const scmf_desc* d, orc_derived* out) {
  Sys s;
  if (!build(d, s)) return -1;
  out->dx = s.dx;
  out->dy = s.dy;
  out->dz = s.dz;
  out->inverse_refbeads = s.inverse_refbeads;
  out->harmonic_normb = s.harmonic_normb;
  out->field_scaling = s.field_scaling_type[0];
  out->A = s.A[0];
  out->R = s.R[0];
  out->n_cells = s.n_cells;
  out->n_chains = s.n_chains;
  out->n_beads = s.n_beads;
  return 0;
}

// One NGP index (device_function.cu:313-360) for tests.
unsigned orc_bin_index(const scmf_desc* d, double rx, double ry, double rz, unsigned type) {
  Sys s;
  if (!build(d, s)) return 0xFFFFFFFFu;
  return bin(s, rx, ry, rz, type);
}

void orc_bin_many(const scmf_desc* d, const double* r, int64_t n, const uint32_t* type,
                  uint32_t* out) {
  Sys s;
  if (!build(d, s)) return;
  for (int64_t i = 0; i < n; ++i) out[i] = bin(s, r[3 * i], r[3 * i + 1], r[3 * i + 2], type[i]);
}

// init_array + coord_to_density, device_function.cu:471-508: zero, then one
// count per bead at its NGP cell.  OpenMP threads keep private histograms and
// sum them (integer counts are order independent).
int orc_grid(const scmf_desc* d, const float* pos, uint32_t* counts) {
  Sys s;
  if (!build(d, s)) return -1;
  const int64_t tc = (int64_t)s.T * s.n_cells;
  memset(counts, 0, sizeof(uint32_t) * tc);
#pragma omp parallel
  {
    std::vector<uint32_t> priv(tc, 0);
#pragma omp for schedule(static)
    for (int64_t chain = 0; chain < s.n_chains; ++chain) {
      const int p = poly_type_of(s, chain);
      const int64_t b0 = first_bead(s, p, chain);
      for (int m = 0; m < s.len[p]; ++m) {
        const int64_t xyz = b0 + m;
        const double x = pos[3 * xyz], y = pos[3 * xyz + 1], z = pos[3 * xyz + 2];
        const unsigned mono_type = (unsigned)s.monotype[s.mono_off[p] + m];
        priv[bin(s, x, y, z, mono_type)] += 1;
      }
    }
#pragma omp critical
    for (int64_t c = 0; c < tc; ++c) counts[c] += priv[c];
  }
  return 0;
}

// omega_field_update, device_function.cu:373-412, fp64, same operation order.
int orc_omega(const scmf_desc* d, const uint32_t* density, double* omega) {
  Sys s;
  if (!build(d, s)) return -1;
  const int64_t C = s.n_cells;
  const unsigned T = (unsigned)s.T;
#pragma omp parallel for schedule(static)
  for (int64_t cell = 0; cell < C; ++cell) {
    for (unsigned Tt = 0; Tt < T; ++Tt) {
      double w = 0;
      for (unsigned S = 0; S < T; ++S) w += (double)density[cell + S * C] * s.field_scaling_type[Tt];
      w += -1;
      w = s.inverse_refbeads * (s.xn[Tt + T * Tt] * w);
      for (unsigned S = 0; S < T; ++S) {
        if (Tt != S) {
          const double dnorm = -0.5 * s.xn[Tt + T * S];
          const double interaction =
              dnorm * s.inverse_refbeads *
              (s.field_scaling_type[Tt] * (double)density[cell + Tt * C] -
               s.field_scaling_type[S] * (double)density[cell + S * C]);
          w += interaction;
        }
      }
      omega[cell + Tt * C] = w;
    }
  }
  return 0;
}

// initialize_coord, device_function.cu:140-207: chain start uniform in the
// box, then a walk over the bond graph placing every unplaced neighbour at
// parent + N(0,1)/sqrt(3(N-1)) per component, N = poly_length.
int orc_init_coords(const scmf_desc* d, const scmf_opts* o, float* pos) {
  Sys s;
  if (!build(d, s)) return -1;
#pragma omp parallel for schedule(static)
  for (int64_t chain = 0; chain < s.n_chains; ++chain) {
    const int p = poly_type_of(s, chain);
    const int64_t pi = first_bead(s, p, chain);
    const int N = s.len[p];
    std::vector<char> occupy(N, 0);
    const u32x4 w0 = stream_block(o->seed, o->system_id, (uint32_t)chain, INIT_SWEEP, 0, SLOT_INIT);
    pos[pi * 3] = (float)((double)u01_24f(w0.v[0]) * s.Lx);
    pos[pi * 3 + 1] = (float)((double)u01_24f(w0.v[1]) * s.Ly);
    pos[pi * 3 + 2] = (float)((double)u01_24f(w0.v[2]) * s.Lz);
    occupy[0] = 1;
    const double reference_Nbeads = N;
    const double harmonic_spring_Cste = 1.0 / sqrt(3.0 * (reference_Nbeads - 1.0));
    for (int i = 0; i < N; i++) {
      const int b = s.mono_off[p] + i;
      for (int k = s.nbr_start[b]; k < s.nbr_start[b + 1]; ++k) {
        const int index = s.nbr[k];
        if (occupy[index] == 0) {
          const u32x4 w = stream_block(o->seed, o->system_id, (uint32_t)chain, INIT_SWEEP,
                                       (uint32_t)(1 + index), SLOT_INIT);
          double nx, ny, nz;
          normal3(w, &nx, &ny, &nz);
          pos[(pi + index) * 3] = (float)(nx * harmonic_spring_Cste + (double)pos[(pi + i) * 3]);
          pos[(pi + index) * 3 + 1] =
              (float)(ny * harmonic_spring_Cste + (double)pos[(pi + i) * 3 + 1]);
          pos[(pi + index) * 3 + 2] =
              (float)(nz * harmonic_spring_Cste + (double)pos[(pi + i) * 3 + 2]);
          occupy[index] = 1;
        }
      }
    }
  }
  return 0;
}

// One sweep (every chain, N moves each) against a fixed fp32 omega field.
// chain_begin/chain_end bound the chains processed (a bounded sample for the
// CPU baseline); pass 0, -1 for all.
int orc_sweep(const scmf_desc* d, const scmf_opts* o, float* pos, const float* omega,
              uint32_t sweep, int64_t chain_begin, int64_t chain_end, uint64_t* n_accept) {
  Sys s;
  if (!build(d, s)) return -1;
  if (chain_end < 0 || chain_end > s.n_chains) chain_end = s.n_chains;
  uint64_t acc = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : acc)
  for (int64_t chain = chain_begin; chain < chain_end; ++chain) {
    PhiloxDraws rnd{*o, sweep, chain, {}, {}};
    acc += chain_sweep(s, *o, pos, omega, chain, rnd);
  }
  if (n_accept) *n_accept = acc;
  return 0;
}

// One sweep driven by injected random numbers (see InjectedDraws), for
// statistical comparisons with generators other than the engine's.
int orc_sweep_injected(const scmf_desc* d, const scmf_opts* o, float* pos, const float* omega,
                       const double* pick, const float* gauss, const float* acc_u,
                       uint64_t* n_accept) {
  Sys s;
  if (!build(d, s)) return -1;
  uint64_t acc = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : acc)
  for (int64_t chain = 0; chain < s.n_chains; ++chain) {
    const int p = poly_type_of(s, chain);
    InjectedDraws rnd{pick, gauss, acc_u};
    rnd.b0 = first_bead(s, p, chain);
    acc += chain_sweep(s, *o, pos, omega, chain, rnd);
  }
  if (n_accept) *n_accept = acc;
  return 0;
}

// test.cpp:45-67 generalised: mean over chains of |r_last - r_first|^2 in fp64.
double orc_re2(const scmf_desc* d, const float* pos) {
  Sys s;
  if (!build(d, s)) return -1;
  double end_dis = 0;
  for (int64_t chain = 0; chain < s.n_chains; ++chain) {
    const int p = poly_type_of(s, chain);
    const int64_t a = first_bead(s, p, chain), b = a + s.len[p] - 1;
    const double x = pos[a * 3], y = pos[a * 3 + 1], z = pos[a * 3 + 2];
    const double x1 = pos[b * 3], y1 = pos[b * 3 + 1], z1 = pos[b * 3 + 2];
    end_dis += (x - x1) * (x - x1) + (y - y1) * (y - y1) + (z - z1) * (z - z1);
  }
  return end_dis / (double)s.n_chains;
}

// Primitive access for the pinning tests.
void orc_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                uint32_t* out) {
  const u32x4 r = philox4x32_10(c0, c1, c2, c3, k0, k1);
  for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}
// Blocks of w (four words each) whose picks of moves 0 and 1 on chains of n
// beads differ between the defining pick_bead(pick_u01(.)) and the engine's
// pick_of (the count of such picks; zero when the two forms agree).
int64_t orc_pick_disagreements(const uint32_t* w, int64_t blocks, int32_t n) {
  const double n53 = (double)n * 0x1p-53;
  int64_t bad = 0;
  for (int64_t k = 0; k < blocks; ++k) {
    u32x4 b;
    for (int j = 0; j < 4; ++j) b.v[j] = w[4 * k + j];
    for (uint32_t i = 0; i < 2; ++i)
      bad += pick_of(b, i, n, n53) != pick_bead(pick_u01(b, i), n) ? 1 : 0;
  }
  return bad;
}
void orc_exp_many(const double* x, double* y, int64_t n) {
  for (int64_t i = 0; i < n; ++i) y[i] = exp_cw(x[i]);
}
void orc_log_many(const double* x, double* y, int64_t n) {
  for (int64_t i = 0; i < n; ++i) y[i] = log_pos(x[i]);
}
void orc_sincos_many(const double* u, double* s, double* c, int64_t n) {
  for (int64_t i = 0; i < n; ++i) sincos_2pi(u[i], &s[i], &c[i]);
}
void orc_normals(uint64_t seed, uint32_t system, uint32_t chain, uint32_t sweep, int64_t n,
                 double* out) {
  for (int64_t i = 0; i < n; ++i) {
    const u32x4 w = stream_block(seed, system, chain, sweep, (uint32_t)i, SLOT_DISP);
    normal3(w, &out[3 * i], &out[3 * i + 1], &out[3 * i + 2]);
  }
}
void orc_logf_many(const float* x, float* y, int64_t n) {
  for (int64_t i = 0; i < n; ++i) y[i] = logf_pos(x[i]);
}
void orc_sincosf_turn24_many(const uint32_t* k, float* s, float* c, int64_t n) {
  for (int64_t i = 0; i < n; ++i) sincosf_turn24(k[i], &s[i], &c[i]);
}
void orc_normals_f(uint64_t seed, uint32_t system, uint32_t chain, uint32_t sweep, int64_t n,
                   float* out) {
  for (int64_t i = 0; i < n; ++i) {
    const u32x4 w = stream_block(seed, system, chain, sweep, (uint32_t)i, SLOT_DISP);
    normal3f(w, &out[3 * i], &out[3 * i + 1], &out[3 * i + 2]);
  }
}
// normal3f of caller-given Philox blocks (4 words each), the host build of
// the device primitive.
void orc_normal3f_words(const uint32_t* w, int64_t n, float* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    u32x4 b;
    for (int k = 0; k < 4; ++k) b.v[k] = w[4 * i + k];
    normal3f(b, &out[3 * i], &out[3 * i + 1], &out[3 * i + 2]);
  }
}
int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
