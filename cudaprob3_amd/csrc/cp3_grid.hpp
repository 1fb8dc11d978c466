// Resident grids shared by the batched engine and the request server: an
// atmospheric grid reduced to its trajectory table on the device, or a beam
// energy list with its density and length.
#pragma once

#include <memory>
#include <vector>

#include "cp3_capi_util.hpp"
#include "cp3_host.hpp"
#include "cp3_launch.hpp"

namespace cp3 {
namespace capi {

struct AtmosGrid {
  int nC = 0, nE = 0, S = 0;
  DevBuf ev, perm, M, Lvac, seg, rye;
  GridDev dev() const {
    GridDev g{};
    g.ev = (const double*)ev.p;
    g.perm = (const int32_t*)perm.p;
    g.M = (const int32_t*)M.p;
    g.Lvac = (const double*)Lvac.p;
    g.seg = (const double*)seg.p;
    g.rye = (const double*)rye.p;
    g.nC = nC;
    g.nE = nE;
    g.S = S;
    return g;
  }
  size_t points() const { return (size_t)nC * (size_t)nE; }
};

struct BeamGrid {
  int nE = 0;
  double rye = 0.0, L = 0.0;
  DevBuf ev;
  size_t points() const { return (size_t)nE; }
};

// Validates the inputs the way setDensity / setCosineList / setEnergyList do
// and uploads the trajectory table.
inline std::unique_ptr<AtmosGrid> make_atmos_grid(const double* cosv, int nC,
                                                  const double* ev, int nE,
                                                  const double* radii,
                                                  const double* rhos,
                                                  const double* yps, int nrows,
                                                  double h_km, int dev) {
  if (nC <= 0 || nE <= 0) throw Error(CP3_ERR_INVALID, "grid sizes must be positive");
  if (!cosv || !ev) throw Error(CP3_ERR_INVALID, "null pointer: cosines or energies");
  if (nrows < 0) throw Error(CP3_ERR_INVALID, "negative number of model rows");
  if (nrows > 0 && (!radii || !rhos || !yps))
    throw Error(CP3_ERR_INVALID, "null pointer: Earth model rows");
  EarthModel em;
  earth_set(em, std::vector<double>(radii, radii + nrows),
            std::vector<double>(rhos, rhos + nrows), std::vector<double>(yps, yps + nrows));
  TrajTable tt;
  build_traj_table(tt, em, cosv, nC, h_km * 100000.0);
  std::unique_ptr<AtmosGrid> g(new AtmosGrid());
  g->nC = nC;
  g->nE = nE;
  g->S = tt.S;
  g->ev.upload(ev, (size_t)nE * sizeof(double), dev);
  g->perm.upload(tt.perm.data(), tt.perm.size() * sizeof(int32_t), dev);
  g->M.upload(tt.M.data(), tt.M.size() * sizeof(int32_t), dev);
  g->Lvac.upload(tt.Lvac.data(), tt.Lvac.size() * sizeof(double), dev);
  g->seg.upload(tt.seg.data(), tt.seg.size() * sizeof(double), dev);
  g->rye.upload(tt.rye.data(), tt.rye.size() * sizeof(double), dev);
  return g;
}

inline std::unique_ptr<BeamGrid> make_beam_grid(const double* ev, int nE, double rho,
                                                double L, int dev) {
  if (nE <= 0) throw Error(CP3_ERR_INVALID, "number of energies must be positive");
  if (!ev) throw Error(CP3_ERR_INVALID, "null pointer: energies");
  std::unique_ptr<BeamGrid> g(new BeamGrid());
  g->nE = nE;
  g->rye = rho * kBeamYe;
  g->L = L;
  g->ev.upload(ev, (size_t)nE * sizeof(double), dev);
  return g;
}

}  // namespace capi
}  // namespace cp3
