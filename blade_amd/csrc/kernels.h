// kernels.h - launchers of the device kernels (one translation unit each).
#pragma once

#include "device.h"

namespace blade {

void launch_domdec_rebuild(const Dev &d, int maxBlocks, hipStream_t st);
void launch_domdec_pack(const Dev &d, int maxBlocks, hipStream_t st);
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
void launch_bonded(const Dev &d, bool energy, bool bias, hipStream_t st);
// the standard normals of the current step (the O step reads them)
void launch_normals(const Dev &d, hipStream_t st);
// a whole update; launches its own normals
void launch_update(const Dev &d, bool energy_step, hipStream_t st);
// needs launch_normals of the same step first
void launch_update_after_kick(const Dev &d, hipStream_t st);
void launch_theta_to_lambda(const Dev &d, hipStream_t st);
void launch_set_step(const Dev &d, long long step, hipStream_t st);

}  // namespace blade
