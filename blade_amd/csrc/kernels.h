// kernels.h - launchers of the device kernels (one translation unit each).
#pragma once

#include "device.h"

namespace blade {

void launch_domdec_rebuild(const Dev &d, int maxBlocks, hipStream_t st);
// the rebuild over several systems: kRebuildKernels launches, one per
// rebuild kernel, each over the systems of its plan
constexpr int kRebuildKernels = 14;
int grid_rebuild(const Dev &d, int k);
void launch_domdec_rebuild_multi(const Multi *m, hipStream_t st);
void launch_domdec_pack(const Dev &d, int maxBlocks, hipStream_t st);
// the pair kernel over the list's chunks (the FP64 dU/dlambda of the
// alchemical atoms is launch_nb_dudl)
void launch_nbdirect(const Dev &d, bool energy, bool alch, bool soft, int nwg, hipStream_t st);
// PME charge grid: bins, brick spread, FFT with the convolution (the
// potential grid is left in rgrid for the force assembly)
void launch_pme_grid(const Dev &d, bool energy, hipStream_t st);
// per-atom force assembly: adds the components in mask (kAsmNb chunk
// partials, kAsmBonded slots, kAsmPme gather) to the force array
// kick = true: also the V and half R of a non-energy step for the atoms
// (launch_update_after_kick does the rest of that step)
void launch_force_assemble(const Dev &d, unsigned mask, hipStream_t st, bool kick = false);
bool pme_grid_supported(int K);
// the term families in fam (kFam*); families left out zero their slots
void launch_bonded(const Dev &d, bool energy, unsigned fam, hipStream_t st);
// the standard normals of the current step (the O step reads them)
void launch_normals(const Dev &d, hipStream_t st);
// a whole update; launches its own normals
void launch_update(const Dev &d, bool energy_step, hipStream_t st);
// needs launch_normals of the same step first
void launch_update_after_kick(const Dev &d, hipStream_t st);
void launch_constrain_positions(const Dev &d, hipStream_t st);
void launch_constrain_velocities(const Dev &d, hipStream_t st);
void launch_chain_rule(const Dev &d, hipStream_t st);
// init: a new theta from the host (with the fix flag it sets the held lambda)
void launch_theta_to_lambda(const Dev &d, hipStream_t st, bool init = false);
// constraint groups at exactly their lengths before the first step
void launch_rectify_init(const Dev &d, hipStream_t st);
// virtual sites: positions from the hosts; the sites' forces onto the hosts
void launch_virtual_place(const Dev &d, hipStream_t st);
void launch_virtual_spread(const Dev &d, hipStream_t st);
// MC barostat: back up and scale every atom, rectify constraint groups
// (oldbox: the box before the move); restore the backup
void launch_volume_move(const Dev &d, const double *scale, const double *oldbox, hipStream_t st);
void launch_volume_restore(const Dev &d, hipStream_t st);
void launch_set_step(const Dev &d, long long step, hipStream_t st);
// the bins and spread of launch_pme_grid alone
void launch_pme_spread(const Dev &d, hipStream_t st);
// the FFT part of launch_pme_grid alone
void launch_pme_fft(const Dev &d, bool energy, hipStream_t st);
// the MFMA split-BF16x3 form of it (every grid edge fft_mfma_supported)
bool fft_mfma_supported(int K);
// the MFMA form over several members with K1 == K2 and equal grids
int grid_fft_mfma_planes(const Dev &d);
int grid_fft_mfma_x(const Dev &d);
void launch_pme_fft_mfma_multi(const Multi &mp, const Multi &mx, int K0, int K1, bool energy, hipStream_t st);
void launch_pme_fft_mfma(const Dev &d, bool energy, hipStream_t st);

// ---- several systems per launch (Multi, device.h) ----------------------
// Workgroups one system needs per kernel type; the single-system launches
// above use the same counts, so a system's work is the same either way.
int grid_pack(const Dev &d);
int grid_normals(const Dev &d);
int grid_bonded(const Dev &d);
int grid_pme_bin(const Dev &d);
int grid_pme_spread(const Dev &d);
size_t lds_pme_spread(const Dev &d);
bool fft_plane_path(const Dev &d);
int grid_fft_planes(const Dev &d);
int grid_fft_x(const Dev &d);
int grid_nbdirect(const Dev &d);
int grid_nb_dudl(const Dev &d);
// FP64 dU/dlambda of the nonbonded pairs, by alchemical atom
void launch_nb_dudl(const Dev &d, bool soft, hipStream_t st);
void launch_nb_dudl_multi(const Multi &m, int elec, int vdw, bool soft, hipStream_t st);
size_t lds_nbdirect(const Dev &d);
int grid_assemble(const Dev &d);
int grid_tail(const Dev &d);
// Every system of a Multi must share the template keys of its kernel: the
// spline order (bin, spread, assembly), the grid sizes (FFT), the methods
// and flags (nbdirect).
void launch_pack_multi(const Multi &m, hipStream_t st);
void launch_normals_multi(const Multi &m, hipStream_t st);
void launch_bonded_multi(const Multi &m, bool energy, hipStream_t st);
void launch_pme_bin_multi(const Multi &m, int order, hipStream_t st);
void launch_pme_spread_multi(const Multi &m, int order, size_t lds, hipStream_t st);
void launch_pme_fft_multi(const Multi &mp, const Multi &mx, int K0, int K1, bool energy, hipStream_t st);
void launch_nbdirect_multi(const Multi &m, int elec, int vdw, bool energy, bool alch, bool soft, size_t lds,
                           hipStream_t st);
void launch_assemble_multi(const Multi &m, int order, unsigned mask, bool kick, hipStream_t st);
void launch_tail_multi(const Multi &m, hipStream_t st);

}  // namespace blade
