// sc_host.cpp -- host layer: reference-style YAML parameter files, derived
// system constants, and filament generation in the reference's "random"
// insertion style.  Pure host C++ (no device code).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <random>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "sc_host.h"

namespace sc_host {

// ---------------------------------------------------------------------------
// YAML subset: block maps, block sequences (of maps or scalars), flow lists,
// plain / quoted scalars, comments.  Enough for the reference's parameter
// files (examples/*.yaml, config/default_config.yaml).
// ---------------------------------------------------------------------------
const YNode *YNode::get(const std::string &k) const {
  if (kind != Map) return nullptr;
  for (const auto &kv : map)
    if (kv.first == k) return &kv.second;
  return nullptr;
}

namespace {

struct Line {
  int indent;
  std::string text;
  int no;
};

std::string rstrip(const std::string &s) {
  size_t e = s.find_last_not_of(" \t\r\n");
  return e == std::string::npos ? std::string() : s.substr(0, e + 1);
}
std::string strip(const std::string &s) {
  size_t b = s.find_first_not_of(" \t\r\n");
  if (b == std::string::npos) return std::string();
  return rstrip(s.substr(b));
}

std::string drop_comment(const std::string &s) {
  char q = 0;
  for (size_t i = 0; i < s.size(); ++i) {
    char c = s[i];
    if (q) {
      if (c == q) q = 0;
    } else if (c == '"' || c == '\'') {
      q = c;
    } else if (c == '#' && (i == 0 || s[i - 1] == ' ' || s[i - 1] == '\t')) {
      return s.substr(0, i);
    }
  }
  return s;
}

std::string unquote(const std::string &s) {
  if (s.size() >= 2 && ((s.front() == '"' && s.back() == '"') ||
                        (s.front() == '\'' && s.back() == '\'')))
    return s.substr(1, s.size() - 2);
  return s;
}

// position of the "key:" separator (a colon followed by space or EOL,
// outside quotes and brackets), or npos
size_t key_sep(const std::string &s) {
  char q = 0;
  int depth = 0;
  for (size_t i = 0; i < s.size(); ++i) {
    char c = s[i];
    if (q) {
      if (c == q) q = 0;
      continue;
    }
    if (c == '"' || c == '\'') q = c;
    else if (c == '[' || c == '{') ++depth;
    else if (c == ']' || c == '}') --depth;
    else if (c == ':' && depth == 0 && (i + 1 == s.size() || s[i + 1] == ' '))
      return i;
  }
  return std::string::npos;
}

YNode scalar_or_flow(const std::string &v, std::string *err, int no) {
  YNode n;
  std::string t = strip(v);
  if (!t.empty() && t.front() == '[') {
    if (t.back() != ']') {
      *err = "line " + std::to_string(no) + ": unterminated flow list";
      return n;
    }
    n.kind = YNode::Seq;
    std::string body = t.substr(1, t.size() - 2);
    std::string cur;
    char q = 0;
    for (char c : body) {
      if (q) {
        cur += c;
        if (c == q) q = 0;
      } else if (c == '"' || c == '\'') {
        q = c;
        cur += c;
      } else if (c == ',') {
        YNode e;
        e.kind = YNode::Scalar;
        e.scalar = unquote(strip(cur));
        n.seq.push_back(e);
        cur.clear();
      } else {
        cur += c;
      }
    }
    if (!strip(cur).empty()) {
      YNode e;
      e.kind = YNode::Scalar;
      e.scalar = unquote(strip(cur));
      n.seq.push_back(e);
    }
    return n;
  }
  n.kind = YNode::Scalar;
  n.scalar = unquote(t);
  return n;
}

struct Parser {
  std::vector<Line> lines;
  size_t i = 0;
  std::string err;

  YNode block(int indent) {
    YNode n;
    if (i >= lines.size()) return n;
    if (lines[i].text.rfind("- ", 0) == 0 || lines[i].text == "-") return seq(indent);
    return map(indent);
  }

  YNode map(int indent) {
    YNode n;
    n.kind = YNode::Map;
    while (i < lines.size() && err.empty()) {
      const Line &L = lines[i];
      if (L.indent < indent) break;
      if (L.indent > indent) {
        err = "line " + std::to_string(L.no) + ": unexpected indentation";
        break;
      }
      if (L.text.rfind("- ", 0) == 0) break;
      size_t sep = key_sep(L.text);
      if (sep == std::string::npos) {
        err = "line " + std::to_string(L.no) + ": expected 'key: value'";
        break;
      }
      std::string key = unquote(strip(L.text.substr(0, sep)));
      std::string val = strip(L.text.substr(sep + 1));
      ++i;
      if (!val.empty()) {
        n.map.emplace_back(key, scalar_or_flow(val, &err, L.no));
      } else if (i < lines.size() && (lines[i].indent > indent ||
                                      (lines[i].indent == indent &&
                                       lines[i].text.rfind("- ", 0) == 0))) {
        n.map.emplace_back(key, block(lines[i].indent));
      } else {
        n.map.emplace_back(key, YNode());
      }
    }
    return n;
  }

  YNode seq(int indent) {
    YNode n;
    n.kind = YNode::Seq;
    while (i < lines.size() && err.empty()) {
      Line &L = lines[i];
      if (L.indent != indent || L.text.rfind("-", 0) != 0) break;
      std::string rest = L.text.size() > 1 ? L.text.substr(2) : std::string();
      std::string rs = strip(rest);
      if (rs.empty()) {
        ++i;
        n.seq.push_back(i < lines.size() && lines[i].indent > indent
                            ? block(lines[i].indent)
                            : YNode());
        continue;
      }
      if (key_sep(rs) != std::string::npos) {
        // "- key: v" opens a map whose content column is indent+2
        const int cindent = indent + 2 + (int)rest.find_first_not_of(' ');
        L.indent = cindent;
        L.text = rs;
        n.seq.push_back(map(cindent));
      } else {
        ++i;
        n.seq.push_back(scalar_or_flow(rs, &err, L.no));
      }
    }
    return n;
  }
};

}  // namespace

bool parse_yaml(const std::string &text, YNode *out, std::string *err) {
  Parser P;
  std::istringstream in(text);
  std::string raw;
  int no = 0;
  while (std::getline(in, raw)) {
    ++no;
    std::string t = rstrip(drop_comment(raw));
    if (strip(t).empty()) continue;
    if (strip(t) == "---") continue;
    size_t ind = t.find_first_not_of(' ');
    if (t.find('\t') != std::string::npos && t.find('\t') < ind) {
      *err = "line " + std::to_string(no) + ": tabs are not allowed for indentation";
      return false;
    }
    P.lines.push_back({(int)ind, t.substr(ind), no});
  }
  if (P.lines.empty()) {
    out->kind = YNode::Map;
    return true;
  }
  *out = P.block(P.lines[0].indent);
  if (P.err.empty() && P.i < P.lines.size())
    P.err = "line " + std::to_string(P.lines[P.i].no) + ": could not parse";
  if (!P.err.empty()) {
    *err = P.err;
    return false;
  }
  return true;
}

// ---------------------------------------------------------------------------
// Defaults of the parameters this path reads (config/default_config.yaml:
// system block lines 4-120, "species" and "filament" blocks).
// ---------------------------------------------------------------------------
namespace {

const std::map<std::string, std::string> &system_defaults() {
  static const std::map<std::string, std::string> d = {
      {"seed", "7859459105545"}, {"n_dim", "3"},           {"n_periodic", "0"},
      {"boundary", "0"},         {"system_radius", "100"}, {"n_steps", "1000000"},
      {"delta", "0.001"},        {"wca_eps", "1"},         {"wca_sig", "1"},
      {"f_cutoff", "2000"},      {"stoch_flag", "true"},   {"potential", "wca"},
      {"like_like_interactions", "true"}, {"interaction_flag", "true"},
      {"run_name", "sc"},        {"n_runs", "1"}};
  return d;
}

const std::map<std::string, std::string> &filament_defaults() {
  static const std::map<std::string, std::string> d = {
      {"num", "0"},
      {"diameter", "1"},
      {"length", "0"},
      {"insertion_type", "random"},
      {"overlap", "false"},
      {"persistence_length", "400"},
      {"max_length", "500"},
      {"min_length", "5"},
      {"min_bond_length", "1.5"},
      {"friction_ratio", "2"},
      {"dynamic_instability_flag", "false"},
      {"polydispersity_flag", "false"},
      {"packing_fraction", "-1"},
      {"driving_factor", "0"},
      {"peclet_number", "-1"},
      {"flexure_number", "-1"},
      {"perlen_ratio", "-1"},
      {"n_equil", "0"},
      {"spiral_flag", "false"},
      {"optical_trap_flag", "false"},
      {"intrinsic_curvature", "0"},
      {"flagella_flag", "false"},
      {"reference_frame_flag", "false"}};
  return d;
}

}  // namespace

std::string Params::sys(const std::string &k) const {
  if (root.kind == YNode::Map)
    if (const YNode *n = root.get(k))
      if (n->kind == YNode::Scalar) return n->scalar;
  auto it = system_defaults().find(k);
  return it == system_defaults().end() ? std::string() : it->second;
}

std::string Params::fil(const std::string &k) const {
  // species block overrides: filament block > shared "species" block >
  // defaults (parse_species_params merges the shared block under the
  // species' own keys)
  if (filament && filament->kind == YNode::Map)
    if (const YNode *n = filament->get(k))
      if (n->kind == YNode::Scalar) return n->scalar;
  if (root.kind == YNode::Map)
    if (const YNode *sp = root.get("species"))
      if (const YNode *n = sp->get(k))
        if (n->kind == YNode::Scalar) return n->scalar;
  auto it = filament_defaults().find(k);
  return it == filament_defaults().end() ? std::string() : it->second;
}

bool load_params(const std::string &path, Params *P, std::string *err) {
  std::ifstream f(path);
  if (!f) {
    *err = "cannot open " + path;
    return false;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  if (!parse_yaml(ss.str(), &P->root, err)) return false;
  P->filament = nullptr;
  if (const YNode *fn = P->root.get("filament")) {
    if (fn->kind == YNode::Map) P->filament = fn;
    else if (fn->kind == YNode::Seq && !fn->seq.empty()) P->filament = &fn->seq[0];
  }
  return true;
}

// Space::InitUnitCell + invert_sym_3d_matrix (src/space.cpp:85-102,
// src/linear_algebra.cpp:52-64), cubic box of side 2R.
void unit_cell(double R, double *h, double *hinv) {
  for (int i = 0; i < 9; ++i) h[i] = 0.0;
  for (int i = 0; i < 3; ++i) h[4 * i] = 2.0 * R;
  const double *a = h;
  double inv_det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                   a[2] * (a[3] * a[7] - a[4] * a[6]);
  inv_det = 1.0 / inv_det;
  double *b = hinv;
  b[0] = inv_det * (a[4] * a[8] - a[5] * a[7]);
  b[4] = inv_det * (a[0] * a[8] - a[2] * a[6]);
  b[8] = inv_det * (a[0] * a[4] - a[1] * a[3]);
  b[1] = b[3] = -inv_det * (a[3] * a[8] - a[5] * a[6]);
  b[2] = b[6] = inv_det * (a[3] * a[7] - a[4] * a[6]);
  b[5] = b[7] = -inv_det * (a[0] * a[7] - a[1] * a[6]);
}

// Filament::InitFilamentLength, static-length branch
// (src/filament.cpp:104-147): number of bonds and bond length.
void filament_layout(double length, double min_bond_length, double min_length,
                     double max_length, int *n_bonds, double *bond_length) {
  if (min_length < 2 * min_bond_length) min_length = 2 * min_bond_length;
  if (max_length < min_length) max_length = min_length;
  if (length > max_length) length = max_length;
  else if (length < min_length) length = min_length;
  double max_bond_length = 1.5 * min_bond_length;
  int nb = (int)std::floor(length / max_bond_length);
  nb = (nb == 1 ? 2 : nb);
  if (nb < 2) nb = 2;
  *n_bonds = nb;
  *bond_length = length / nb;
}

// Filament::SetDiffusion (src/filament.cpp:320-348)
void friction(double length, double diameter, int n_sites, double friction_ratio,
              double *gamma_par, double *gamma_perp) {
  double logLD = std::log(length / diameter);
  double gp = 4.0 * length / (3.0 * n_sites * logLD);
  *gamma_perp = gp;
  *gamma_par = gp / friction_ratio;
}

// ---------------------------------------------------------------------------
// Random filament insertion.
// ---------------------------------------------------------------------------
namespace {

struct Rng {
  std::mt19937_64 g;
  explicit Rng(uint64_t s) : g(s) {}
  double uniform_pos() {  // (0,1)
    for (;;) {
      double u = (double)(g() >> 11) * (1.0 / 9007199254740992.0);
      if (u > 0.0) return u;
    }
  }
};

// RNG::RandomUnitVector, 3D (src/rng.cpp:30-45)
void random_unit_vector(Rng &rng, double *v) {
  double z = 2.0 * rng.uniform_pos() - 1.0;
  double w = std::sqrt(1 - z * z);
  double t = 2.0 * M_PI * rng.uniform_pos();
  v[0] = w * std::cos(t);
  v[1] = w * std::sin(t);
  v[2] = z;
}

// rotate_vector_relative, 3D (src/linear_algebra.cpp:230-251)
void rotate_relative(double *v1, const double *v2) {
  double vx = v1[0], vy = v1[1], vz = v1[2];
  double phi = std::atan2(v2[1], v2[0]);
  double theta = std::acos(v2[2]);
  v1[0] = vx * std::cos(theta) * std::cos(phi) - vy * std::sin(phi) +
          vz * std::cos(phi) * std::sin(theta);
  v1[1] = vx * std::cos(theta) * std::sin(phi) + vy * std::cos(phi) +
          vz * std::sin(theta) * std::sin(phi);
  v1[2] = vz * std::cos(theta) - vx * std::sin(theta);
}

// Filament::GenerateProbableOrientation, 3D (src/filament.cpp:350-383)
void probable_orientation(Rng &rng, double kappa, double bl, double *u) {
  double theta;
  double k = kappa / bl;
  if (kappa == 0) {
    theta = rng.uniform_pos() * M_PI;
  } else if (kappa < 100) {
    theta = std::acos(std::log(std::exp(-k) + 2.0 * rng.uniform_pos() * std::sinh(k)) / k);
  } else {
    theta = std::acos((std::log(2.0 * rng.uniform_pos()) - std::log(2.0) + k) / k);
  }
  double phi = rng.uniform_pos() * 2.0 * M_PI;
  double nu[3] = {std::sin(theta) * std::cos(phi), std::sin(theta) * std::sin(phi),
                  std::cos(theta)};
  rotate_relative(nu, u);
  u[0] = nu[0];
  u[1] = nu[1];
  u[2] = nu[2];
}

// Closest distance^2 between two segments (centre c, unit u, half length h)
// under the periodic minimum image of a cubic box of side L.
double seg_dist2(const double *c1, const double *u1, double h1, const double *c2,
                 const double *u2, double h2, double L) {
  double d[3];
  for (int k = 0; k < 3; ++k) {
    d[k] = c2[k] - c1[k];
    d[k] -= L * std::round(d[k] / L);
  }
  double a = d[0] * u1[0] + d[1] * u1[1] + d[2] * u1[2];
  double b = d[0] * u2[0] + d[1] * u2[1] + d[2] * u2[2];
  double c = u1[0] * u2[0] + u1[1] * u2[1] + u1[2] * u2[2];
  double den = 1.0 - c * c;
  double s = 0, t = 0;
  if (den > 1e-12) {
    s = (a - c * b) / den;
  }
  auto clamp = [](double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); };
  s = clamp(s, -h1, h1);
  t = clamp(c * s - b, -h2, h2);
  s = clamp(a + c * t, -h1, h1);
  t = clamp(c * s - b, -h2, h2);
  double m2 = 0;
  for (int k = 0; k < 3; ++k) {
    double r = d[k] + t * u2[k] - s * u1[k];
    m2 += r * r;
  }
  return m2;
}

// Dense uniform grid of accepted bond centres (head/next linked lists).
struct Grid {
  double L, cl;
  int n;
  std::vector<int32_t> head, next;
  void init(double L_, double min_cell) {
    L = L_;
    n = std::max(1, (int)std::floor(L / min_cell));
    if ((int64_t)n * n * n > (int64_t)64 << 20) n = 400;  // bounded memory
    cl = L / n;
    head.assign((size_t)n * n * n, -1);
  }
  int idx(double x) const {
    double s = x / L - std::floor(x / L);
    int c = (int)(s * n);
    return c >= n ? n - 1 : c;
  }
  int64_t key(int x, int y, int z) const {
    x = ((x % n) + n) % n;
    y = ((y % n) + n) % n;
    z = ((z % n) + n) % n;
    return ((int64_t)x * n + y) * n + z;
  }
  void add(int j, const double *c) {
    int64_t k = key(idx(c[0]), idx(c[1]), idx(c[2]));
    if ((int)next.size() <= j) next.resize(j + 1, -1);
    next[j] = head[k];
    head[k] = j;
  }
};

}  // namespace

int generate(double R, int nf, const int *n_sites, const double *length,
             const double *kappa, const double *diameter, double friction_ratio,
             uint64_t seed, std::vector<int32_t> *site_off, std::vector<double> *r,
             std::vector<sc_filament_params> *fil, std::string *err) {
  const double L = 2.0 * R;
  double maxd = 0, maxbl = 0;
  for (int f = 0; f < nf; ++f) {
    if (n_sites[f] < 3) {
      *err = "filaments need at least 3 sites";
      return -1;
    }
    maxd = std::max(maxd, diameter[f]);
    maxbl = std::max(maxbl, length[f] / (n_sites[f] - 1));
  }
  Grid G;
  G.init(L, maxd + maxbl);
  // Long filaments are placed first (into the emptier box); the output
  // keeps the caller's filament order.
  std::vector<int> order(nf);
  for (int f = 0; f < nf; ++f) order[f] = f;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return n_sites[a] > n_sites[b]; });
  std::vector<std::vector<double>> placed_pts(nf);
  std::vector<double> bc, bu, bh, bd;  // accepted bonds
  fil->assign(nf, sc_filament_params{});
  Rng rng(seed);
  int64_t total_attempts = 0;
  for (int f : order) {
    const int ns = n_sites[f];
    const double bl = length[f] / (ns - 1);
    const double d = diameter[f];
    bool placed = false;
    std::vector<double> pts(3 * (size_t)ns), own(6 * (size_t)ns);
    const int self_skip = 2 + (int)std::ceil(d / bl);
    int64_t attempts_f = 0;
    for (int attempt = 0; attempt < 1000000 && !placed; ++attempt) {
      ++attempts_f;
      // Mesh::InitRandomBond + Filament::InsertFilament
      // (src/mesh.cpp:244-258, src/filament.cpp:225-249), each new bond
      // tested against the accepted bonds as it is grown
      // (InteractionManager::CheckOverlap: dr^2 < (d1+d2)^2/4); an attempt
      // stops at its first overlap.
      double u[3], pos[3];
      random_unit_vector(rng, u);
      for (int k = 0; k < 3; ++k) pos[k] = (2.0 * rng.uniform_pos() - 1.0) * (R - d);
      for (int k = 0; k < 3; ++k) pos[k] -= 0.5 * u[k] * length[f];
      for (int k = 0; k < 3; ++k) pts[k] = pos[k];
      bool ok = true;
      for (int i = 1; i < ns && ok; ++i) {
        if (i >= 2) probable_orientation(rng, kappa[f], bl, u);
        for (int k = 0; k < 3; ++k) pts[3 * i + k] = pts[3 * (i - 1) + k] + bl * u[k];
        double c[3], uu[3], l2 = 0;
        for (int k = 0; k < 3; ++k) {
          uu[k] = pts[3 * i + k] - pts[3 * (i - 1) + k];
          l2 += uu[k] * uu[k];
        }
        double l = std::sqrt(l2);
        for (int k = 0; k < 3; ++k) {
          c[k] = pts[3 * (i - 1) + k] + 0.5 * uu[k];
          uu[k] /= l;
        }
        // the filament's own earlier bonds more than a diameter of contour
        // away (a long filament in a small periodic box can cross its own
        // images)
        for (int j = 0; j + self_skip < i && ok; ++j) {
          const double *cj = &own[6 * j];
          if (seg_dist2(c, uu, 0.5 * l, cj, cj + 3, 0.5 * l, L) < d * d) ok = false;
        }
        for (int k = 0; k < 3; ++k) {
          own[6 * (i - 1) + k] = c[k];
          own[6 * (i - 1) + 3 + k] = uu[k];
        }
        int cx = G.idx(c[0]), cy = G.idx(c[1]), cz = G.idx(c[2]);
        for (int dx = -1; dx <= 1 && ok; ++dx)
          for (int dy = -1; dy <= 1 && ok; ++dy)
            for (int dz = -1; dz <= 1 && ok; ++dz) {
              for (int j = G.head[G.key(cx + dx, cy + dy, cz + dz)]; j >= 0 && ok;
                   j = G.next[j]) {
                double lim = 0.5 * (d + bd[j]);
                if (seg_dist2(c, uu, 0.5 * l, &bc[3 * j], &bu[3 * j], bh[j], L) < lim * lim)
                  ok = false;
              }
            }
      }
      if (!ok) continue;
      placed = true;
      for (int i = 0; i < ns - 1; ++i) {
        double c[3], uu[3], l2 = 0;
        for (int k = 0; k < 3; ++k) {
          uu[k] = pts[3 * (i + 1) + k] - pts[3 * i + k];
          l2 += uu[k] * uu[k];
        }
        double l = std::sqrt(l2);
        for (int k = 0; k < 3; ++k) {
          c[k] = pts[3 * i + k] + 0.5 * uu[k];
          uu[k] /= l;
        }
        int j = (int)bh.size();
        for (int k = 0; k < 3; ++k) {
          bc.push_back(c[k]);
          bu.push_back(uu[k]);
        }
        bh.push_back(0.5 * l);
        bd.push_back(d);
        G.add(j, c);
      }
    }
    if (!placed) {
      *err = "insertion failed for filament " + std::to_string(f) + " (" +
             std::to_string(ns) + " sites) after " + std::to_string(attempts_f) + " attempts";
      return -1;
    }
    total_attempts += attempts_f;
    placed_pts[f] = pts;
    sc_filament_params q{};
    q.bond_length = bl;
    q.persistence_length = kappa[f];
    q.diameter = d;
    friction(length[f], d, ns, friction_ratio, &q.gamma_par, &q.gamma_perp);
    (*fil)[f] = q;
  }
  site_off->assign(1, 0);
  std::vector<double> xyz;
  for (int f = 0; f < nf; ++f) {
    xyz.insert(xyz.end(), placed_pts[f].begin(), placed_pts[f].end());
    site_off->push_back(site_off->back() + n_sites[f]);
  }
  if (std::getenv("SIMCORE_AMD_VERBOSE"))
    std::fprintf(stderr, "simcore_amd: inserted %d filaments in %lld attempts\n", nf,
                 (long long)total_attempts);
  const int64_t NS = site_off->back();
  r->assign(3 * (size_t)NS, 0.0);
  for (int64_t i = 0; i < NS; ++i)
    for (int k = 0; k < 3; ++k) (*r)[k * NS + i] = xyz[3 * i + k];
  return 0;
}

}  // namespace sc_host

// ---------------------------------------------------------------------------
// C ABI of the host layer
// ---------------------------------------------------------------------------
extern "C" {

int sc_params_from_yaml(const char *path, sc_params *out, double *spec) {
  if (!path || !out) return SC_ERR_ARG;
  sc_host::Params P;
  std::string err;
  if (!sc_host::load_params(path, &P, &err)) {
    std::fprintf(stderr, "simcore_amd: %s\n", err.c_str());
    return SC_ERR_PARSE;
  }
  sc_params p{};
  try {
    p.n_dim = std::stoi(P.sys("n_dim"));
    p.n_periodic = std::stoi(P.sys("n_periodic"));
    if (p.n_dim != 3 || p.n_periodic != 3 || std::stoi(P.sys("boundary")) != 0) return SC_ERR_ARG;
    if (P.sys("potential") != "wca") return SC_ERR_ARG;
    p.system_radius = std::stod(P.sys("system_radius"));
    p.delta = std::stod(P.sys("delta"));
    p.wca_eps = std::stod(P.sys("wca_eps"));
    p.wca_sig = std::stod(P.sys("wca_sig"));
    p.f_cutoff = std::stod(P.sys("f_cutoff"));
    p.stoch_flag = sc_host::to_bool_str(P.sys("stoch_flag"));
    p.seed = (uint64_t)std::stoull(P.sys("seed"));
    sc_host::unit_cell(p.system_radius, p.unit_cell, p.unit_cell_inv);
    double mbl = std::stod(P.fil("min_bond_length"));
    double d = std::stod(P.fil("diameter"));
    bool di = sc_host::to_bool_str(P.fil("dynamic_instability_flag"));
    // Simulation::InitSpecies: SetMinCellLength(max(spec_length, d))
    // with spec_length from FilamentSpecies::GetSpecLength
    // (src/simulation.cpp:245-251, src/filament_species.cpp:155-161)
    double spec_len = di ? 2 * mbl : 1.5 * mbl;
    int num = std::stoi(P.fil("num"));
    p.min_cell_length = num > 0 ? (d > spec_len ? d : spec_len) : -1;
    p.precision = SC_PREC_FP64;
    p.use_graph = 0;
    if (spec) {
      spec[0] = num;
      spec[1] = std::stod(P.fil("length"));
      spec[2] = std::stod(P.fil("persistence_length"));
      spec[3] = d;
      spec[4] = mbl;
      spec[5] = std::stod(P.fil("friction_ratio"));
      spec[6] = std::stod(P.sys("n_steps"));
      spec[7] = std::stod(P.fil("min_length"));
      spec[8] = std::stod(P.fil("max_length"));
    }
  } catch (...) {
    return SC_ERR_PARSE;
  }
  *out = p;
  return SC_OK;
}

int sc_params_init(sc_params *out, double system_radius) {
  if (!out || !(system_radius > 0)) return SC_ERR_ARG;
  sc_params p{};
  p.n_dim = 3;
  p.n_periodic = 3;
  p.system_radius = system_radius;
  sc_host::unit_cell(system_radius, p.unit_cell, p.unit_cell_inv);
  p.min_cell_length = -1;
  p.delta = 0.001;
  p.wca_eps = 1;
  p.wca_sig = 1;
  p.f_cutoff = 2000;
  p.stoch_flag = 1;
  p.precision = SC_PREC_FP64;
  p.use_graph = 0;
  p.seed = 7859459105545ull;
  p.capacity_factor = 0;
  *out = p;
  return SC_OK;
}

int sc_filament_layout(double length, double min_bond_length, double min_length,
                       double max_length, int32_t *n_bonds, double *bond_length) {
  if (!n_bonds || !bond_length || !(min_bond_length > 0)) return SC_ERR_ARG;
  int nb;
  sc_host::filament_layout(length, min_bond_length, min_length, max_length, &nb, bond_length);
  *n_bonds = nb;
  return SC_OK;
}

int sc_generate_crosslinks(const sc_params *params, int32_t nf, const int32_t *site_off,
                           const double *r, const sc_filament_params *fil, int64_t nx,
                           double r_bind, uint64_t seed, int32_t *bonds, double *lambda) {
  if (!params || nf <= 0 || !site_off || !r || !fil || nx < 0 || !(r_bind > 0))
    return SC_ERR_ARG;
  if (nx > 0 && (!bonds || !lambda)) return SC_ERR_ARG;
  const double L = 2.0 * params->system_radius;
  const int64_t ns = site_off[nf], nb = ns - nf;
  if (nb <= 1) return SC_ERR_ARG;
  // bond centres (wrapped into the box), orientations, half lengths
  std::vector<double> c(3 * nb), u(3 * nb), hl(nb);
  std::vector<int32_t> bf(nb);
  double lmax = 0;
  for (int f = 0; f < nf; ++f)
    for (int64_t i = site_off[f]; i < site_off[f + 1] - 1; ++i) {
      const int64_t b = i - f;
      double o[3], l = 0;
      for (int k = 0; k < 3; ++k) {
        o[k] = r[k * ns + i + 1] - r[k * ns + i];
        l += o[k] * o[k];
      }
      l = std::sqrt(l);
      for (int k = 0; k < 3; ++k) {
        double x = r[k * ns + i] + 0.5 * o[k];
        x -= L * std::floor(x / L + 0.5);
        c[3 * b + k] = x;
        u[3 * b + k] = o[k] / l;
      }
      hl[b] = 0.5 * l;
      bf[b] = f;
      lmax = std::max(lmax, l);
    }
  // grid of bond centres, cells no smaller than the reach
  const double reach = r_bind + 0.5 * lmax;
  int n = (int)std::floor(L / reach);
  if (n < 1) n = 1;
  if ((int64_t)n * n * n > ((int64_t)1 << 24)) n = 256;
  std::vector<int32_t> head((size_t)n * n * n, -1), next(nb, -1);
  auto idx = [&](double x) {
    int i = (int)std::floor((x / L + 0.5) * n);
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
  };
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t k = ((int64_t)idx(c[3 * b]) * n + idx(c[3 * b + 1])) * n + idx(c[3 * b + 2]);
    next[b] = head[k];
    head[k] = (int32_t)b;
  }
  sc_host::Rng rng(seed);
  std::vector<int32_t> cand;
  for (int64_t x = 0; x < nx; ++x) {
    bool placed = false;
    for (int attempt = 0; attempt < 1000 && !placed; ++attempt) {
      const int32_t b1 = (int32_t)std::min<int64_t>(nb - 1, (int64_t)(rng.uniform_pos() * nb));
      const double lam1 = rng.uniform_pos() * fil[bf[b1]].bond_length;
      double P[3];
      for (int k = 0; k < 3; ++k) P[k] = c[3 * b1 + k] - (hl[b1] - lam1) * u[3 * b1 + k];
      cand.clear();
      const int span = n < 3 ? n : 3, off = n < 3 ? 0 : -1;
      const int cx = idx(P[0] - L * std::floor(P[0] / L + 0.5));
      const int cy = idx(P[1] - L * std::floor(P[1] / L + 0.5));
      const int cz = idx(P[2] - L * std::floor(P[2] / L + 0.5));
      for (int dx = 0; dx < span; ++dx)
        for (int dy = 0; dy < span; ++dy)
          for (int dz = 0; dz < span; ++dz) {
            const int X = ((cx + off + dx) % n + n) % n, Y = ((cy + off + dy) % n + n) % n;
            const int Z = ((cz + off + dz) % n + n) % n;
            for (int32_t b = head[((int64_t)X * n + Y) * n + Z]; b >= 0; b = next[b]) {
              if (b == b1) continue;
              double d[3], t = 0, d2 = 0;
              for (int k = 0; k < 3; ++k) {
                d[k] = c[3 * b + k] - P[k];
                d[k] -= L * std::round(d[k] / L);
                t += d[k] * u[3 * b + k];
              }
              t = std::max(-hl[b], std::min(hl[b], t));
              for (int k = 0; k < 3; ++k) {
                const double e = d[k] - t * u[3 * b + k];
                d2 += e * e;
              }
              if (d2 <= r_bind * r_bind) cand.push_back(b);
            }
          }
      if (cand.empty()) continue;
      std::sort(cand.begin(), cand.end());
      const int32_t b2 = cand[std::min<size_t>(cand.size() - 1,
                                               (size_t)(rng.uniform_pos() * cand.size()))];
      bonds[2 * x] = b1;
      bonds[2 * x + 1] = b2;
      lambda[2 * x] = lam1;
      lambda[2 * x + 1] = rng.uniform_pos() * fil[bf[b2]].bond_length;
      placed = true;
    }
    if (!placed) {
      std::fprintf(stderr, "simcore_amd: crosslink %lld found no bond within range\n",
                   (long long)x);
      return SC_ERR_ARG;
    }
  }
  return SC_OK;
}

int sc_generate_filaments(const sc_params *params, int32_t nf, const int32_t *n_sites,
                          const double *length, const double *persistence_length,
                          const double *diameter, double friction_ratio, uint64_t seed,
                          int32_t *site_off, double *r, sc_filament_params *fil,
                          int64_t *n_sites_total) {
  if (!params || nf <= 0 || !n_sites || !length || !persistence_length || !diameter)
    return SC_ERR_ARG;
  int64_t tot = 0;
  for (int f = 0; f < nf; ++f) tot += n_sites[f];
  if (n_sites_total) *n_sites_total = tot;
  if (!r) return SC_OK;
  if (!site_off || !fil) return SC_ERR_ARG;
  std::vector<int32_t> so;
  std::vector<double> rr;
  std::vector<sc_filament_params> fp;
  std::string err;
  std::vector<int> ns(n_sites, n_sites + nf);
  if (sc_host::generate(params->system_radius, nf, ns.data(), length, persistence_length,
                        diameter, friction_ratio > 0 ? friction_ratio : 2.0, seed, &so, &rr,
                        &fp, &err) != 0) {
    std::fprintf(stderr, "simcore_amd: %s\n", err.c_str());
    return SC_ERR_ARG;
  }
  std::memcpy(site_off, so.data(), sizeof(int32_t) * so.size());
  std::memcpy(r, rr.data(), sizeof(double) * rr.size());
  std::memcpy(fil, fp.data(), sizeof(sc_filament_params) * fp.size());
  return SC_OK;
}

}  // extern "C"
