// sc_output.cpp -- the reference's binary trajectory files for filaments,
// written from caller arrays (host only, no device).
//
// Layouts (native endianness, no padding), as the reference writes them:
//   header (SpeciesBase::InitPositFile / InitSpecFile,
//           src/species_base.cpp:13-27, 55-70):
//     int n_steps, int n_posit (or n_spec), double delta
//   frame (Species::WritePosits / WriteSpecs,
//          include/simcore/species.hpp:221-236): int n_members, then
//   per filament, spec (Mesh::WriteSpec + Filament::WriteSpec,
//          src/mesh.cpp:482-492, src/filament.cpp:1303-1318):
//     double diameter, double length, double bond_length, int n_sites,
//     n_sites x double[3] site positions, double bending_stiffness,
//     unsigned char polymerization state (0 = grow)
//   per filament, posit (Filament::WritePosit, src/filament.cpp:1338-1352):
//     double[3] mean site position, double[3] its scaled position wrapped
//     into [-1/2, 1/2) (Object::UpdatePeriodic), double[3] normalised sum of
//     site orientations (Mesh::GetAvgOrientation, src/mesh.cpp:598-608),
//     double diameter, double length
//   thermo file (OutputManagerBase::InitThermo / WriteThermo,
//          include/simcore/output_manager.hpp:156-183): header int n_steps,
//     int n_thermo, double delta, int n_dim; per frame double[9] unit cell,
//     double[9] pressure tensor, double pressure, double volume
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/simcore_amd.h"

namespace {

struct File {
  FILE *f = nullptr;
  explicit File(const char *path, const char *mode) { f = std::fopen(path, mode); }
  ~File() {
    if (f) std::fclose(f);
  }
  template <typename T>
  bool put(const T &v) {
    return std::fwrite(&v, sizeof(T), 1, f) == 1;
  }
};

bool valid(int32_t nf, const int32_t *so, const double *r, const sc_filament_params *fil) {
  if (nf < 0 || !so || (nf > 0 && (!r || !fil)) || so[0] != 0) return false;
  for (int f = 0; f < nf; ++f)
    if (so[f + 1] - so[f] < 2) return false;
  return true;
}

// nearest integer as the reference's NINT (include/simcore/macros.hpp:6)
inline double nint(double x) { return x < 0.0 ? (double)(int)(x - 0.5) : (double)(int)(x + 0.5); }

}  // namespace

extern "C" {

int sc_output_header(const char *path, int32_t n_steps, int32_t n_frames_every, double delta) {
  if (!path) return SC_ERR_ARG;
  File o(path, "wb");
  if (!o.f) return SC_ERR_ARG;
  if (!o.put(n_steps) || !o.put(n_frames_every) || !o.put(delta)) return SC_ERR_ARG;
  return SC_OK;
}

int sc_spec_frame(const char *path, int32_t nf, const int32_t *site_off, const double *r,
                  const sc_filament_params *fil) {
  if (!path || !valid(nf, site_off, r, fil)) return SC_ERR_ARG;
  File o(path, "ab");
  if (!o.f) return SC_ERR_ARG;
  const int64_t ns = site_off[nf];
  bool ok = o.put(nf);
  const unsigned char poly = 0;
  for (int f = 0; f < nf && ok; ++f) {
    const int32_t n = site_off[f + 1] - site_off[f];
    const double bl = fil[f].bond_length, length = bl * (n - 1);
    ok = o.put(fil[f].diameter) && o.put(length) && o.put(bl) && o.put(n);
    for (int32_t i = 0; i < n && ok; ++i)
      for (int k = 0; k < 3 && ok; ++k) ok = o.put(r[k * ns + site_off[f] + i]);
    ok = ok && o.put(fil[f].persistence_length) && o.put(poly);
  }
  return ok ? SC_OK : SC_ERR_ARG;
}

int sc_posit_frame(const char *path, const sc_params *params, int32_t nf,
                   const int32_t *site_off, const double *r, const sc_filament_params *fil) {
  if (!path || !params || !valid(nf, site_off, r, fil)) return SC_ERR_ARG;
  File o(path, "ab");
  if (!o.f) return SC_ERR_ARG;
  const int64_t ns = site_off[nf];
  bool ok = o.put(nf);
  for (int f = 0; f < nf && ok; ++f) {
    const int64_t s0 = site_off[f], n = site_off[f + 1] - s0;
    double p[3] = {0, 0, 0}, u[3] = {0, 0, 0};
    for (int64_t i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) p[k] += r[k * ns + s0 + i];
    for (int k = 0; k < 3; ++k) p[k] /= (double)n;
    // site orientations: bond i for site i, the last bond for the last site
    for (int64_t i = 0; i < n; ++i) {
      const int64_t j = i < n - 1 ? i : n - 2;
      double d[3], m = 0;
      for (int k = 0; k < 3; ++k) {
        d[k] = r[k * ns + s0 + j + 1] - r[k * ns + s0 + j];
        m += d[k] * d[k];
      }
      m = std::sqrt(m);
      for (int k = 0; k < 3; ++k) u[k] += d[k] / m;
    }
    double un = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    for (int k = 0; k < 3; ++k) u[k] /= un;
    double s[3];
    for (int i = 0; i < 3; ++i) {
      s[i] = 0.0;
      for (int j = 0; j < 3; ++j) s[i] += params->unit_cell_inv[3 * i + j] * p[j];
      s[i] -= nint(s[i]);
    }
    const double length = fil[f].bond_length * (double)(n - 1);
    for (int k = 0; k < 3 && ok; ++k) ok = o.put(p[k]);
    for (int k = 0; k < 3 && ok; ++k) ok = o.put(s[k]);
    for (int k = 0; k < 3 && ok; ++k) ok = o.put(u[k]);
    ok = ok && o.put(fil[f].diameter) && o.put(length);
  }
  return ok ? SC_OK : SC_ERR_ARG;
}

int sc_thermo_header(const char *path, int32_t n_steps, int32_t n_thermo, double delta,
                     int32_t n_dim) {
  if (!path) return SC_ERR_ARG;
  File o(path, "wb");
  if (!o.f) return SC_ERR_ARG;
  if (!o.put(n_steps) || !o.put(n_thermo) || !o.put(delta) || !o.put(n_dim)) return SC_ERR_ARG;
  return SC_OK;
}

int sc_thermo_frame(const char *path, const double *unit_cell, const double *pressure_tensor,
                    double pressure, double volume) {
  if (!path || !unit_cell || !pressure_tensor) return SC_ERR_ARG;
  File o(path, "ab");
  if (!o.f) return SC_ERR_ARG;
  bool ok = true;
  for (int i = 0; i < 9 && ok; ++i) ok = o.put(unit_cell[i]);
  for (int i = 0; i < 9 && ok; ++i) ok = o.put(pressure_tensor[i]);
  ok = ok && o.put(pressure) && o.put(volume);
  return ok ? SC_OK : SC_ERR_ARG;
}

}  // extern "C"
