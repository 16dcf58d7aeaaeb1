// sc_host.h -- host layer of libsimcore_amd: YAML parameter files and
// filament generation (no device code).
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/simcore_amd.h"

namespace sc_host {

struct YNode {
  enum Kind { Null, Scalar, Map, Seq } kind = Null;
  std::string scalar;
  std::vector<std::pair<std::string, YNode>> map;
  std::vector<YNode> seq;
  const YNode *get(const std::string &k) const;
};

bool parse_yaml(const std::string &text, YNode *out, std::string *err);

// A parameter file merged over the reference defaults.
struct Params {
  YNode root;
  const YNode *filament = nullptr;  // first filament species block
  std::string sys(const std::string &key) const;
  std::string fil(const std::string &key) const;
};

bool load_params(const std::string &path, Params *P, std::string *err);
inline bool to_bool_str(const std::string &s) {
  return s == "true" || s == "True" || s == "TRUE" || s == "1" || s == "yes" || s == "on";
}
void unit_cell(double R, double *h, double *hinv);
void filament_layout(double length, double min_bond_length, double min_length,
                     double max_length, int *n_bonds, double *bond_length);
void friction(double length, double diameter, int n_sites, double friction_ratio,
              double *gamma_par, double *gamma_perp);
int generate(double R, int nf, const int *n_sites, const double *length,
             const double *kappa, const double *diameter, double friction_ratio,
             uint64_t seed, std::vector<int32_t> *site_off, std::vector<double> *r,
             std::vector<sc_filament_params> *fil, std::string *err);

}  // namespace sc_host
