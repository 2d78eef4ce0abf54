// batch.h -- the mdg_batch handle of include/mdgpu.h (shared by api.cpp
// and record.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <vector>

#include "engine.h"
#include "mdgpu.h"

struct mdg_batch {
  int precision = MDG_FP64;
  int real_mode = MDG_TABLE;
  std::vector<mdg::Box*> boxes;
  std::vector<std::array<int, 3>> cells;  // unit cells nx, ny, nz per box (files/out)
  std::vector<int> mol_unit;              // ions per formula unit (files/out per mole)
  mdg::BoxSet* set = nullptr;
  hipStream_t stream = nullptr;
};

// mdg_scalars from a box's device scalar array (kScal* layout)
inline void fill_scalars(const double* S, mdg_scalars* out) {
  using namespace mdg;
  out->pot = S[kScalPot];
  out->pot_real = S[kScalPotReal];
  out->pot_recip = S[kScalPotRecip];
  out->pot_self = S[kScalPotSelf];
  out->etot = S[kScalEtot];
  out->vir[0] = S[kScalVirX]; out->vir[1] = S[kScalVirY]; out->vir[2] = S[kScalVirZ];
  out->kin = S[kScalKin];
  out->kin_xyz[0] = S[kScalKinX]; out->kin_xyz[1] = S[kScalKinY]; out->kin_xyz[2] = S[kScalKinZ];
  out->pres = S[kScalPres];
  out->pres_xyz[0] = S[kScalPresX]; out->pres_xyz[1] = S[kScalPresY];
  out->pres_xyz[2] = S[kScalPresZ];
  out->step = (long long)S[kScalStep];
}
