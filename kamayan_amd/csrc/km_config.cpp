// km_config.cpp -- option surface of the hydro unit: the reference's input-deck keys,
// defaults and allowed values, parsed into km_cfg_t.
//
// Keys and defaults follow hydro::SetupParams (hydro.cpp:52-91), physics::SetupParams
// (physics.cpp:18-24), driver::SetupParams (kamayan_driver.cpp:23-81), eos
// SetupSpeciesParams (eos.cpp:32-40) and grid::SetupParams (grid.cpp:29-56). An
// invalid value for a known key is rejected, as the reference's runtime-parameter
// rules do (runtime_parameters.hpp:157-169).
#include <cctype>
#include <cfloat>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "../../include/km_hydro.h"
#include "km_internal.hpp"

namespace {

std::string trim(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
  while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
  return s.substr(a, b - a);
}

bool parse_real(const std::string &v, double *out) {
  char *end = nullptr;
  const double x = std::strtod(v.c_str(), &end);
  if (end == v.c_str() || *end != '\0') return false;
  *out = x;
  return true;
}

bool ends_with(const std::string &s, const char *suffix) {
  const size_t n = std::strlen(suffix);
  return s.size() >= n && s.compare(s.size() - n, n, suffix) == 0;
}

bool parse_int(const std::string &v, long *out) {
  char *end = nullptr;
  const long x = std::strtol(v.c_str(), &end, 10);
  if (end == v.c_str() || *end != '\0') return false;
  *out = x;
  return true;
}

template <size_t N>
int pick(const std::string &key, const std::string &val, const char *const (&names)[N],
         const int (&values)[N], int32_t *out) {
  for (size_t i = 0; i < N; ++i)
    if (val == names[i]) {
      *out = values[i];
      return KM_OK;
    }
  km::set_error("invalid value '" + val + "' for " + key);
  return KM_EINVAL;
}

}  // namespace

extern "C" {

int km_cfg_default(km_cfg_t *cfg) {
  if (!cfg) return KM_EINVAL;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->recon = KM_RECON_FOG;           // hydro.cpp:55 default "fog"
  cfg->limiter = KM_LIM_MINMOD;        // hydro.cpp:63 default "minmod"
  cfg->riemann = KM_RIEMANN_HLL;       // hydro.cpp:68 default "hll"
  cfg->recon_vars = KM_RECONVARS_PRIMITIVE;
  cfg->strategy = KM_STRATEGY_SCRATCHPAD;
  cfg->emf_averaging = KM_EMF_ARITHMETIC;
  cfg->fluid = KM_FLUID_ONET;
  cfg->mhd = KM_MHD_OFF;
  cfg->geometry = KM_GEOM_CARTESIAN;
  cfg->nstages = 2;                    // kamayan_driver.cpp:25 default "rk2"
  cfg->precision = KM_PREC_FP64;
  cfg->ndim = 3;
  cfg->gamma = 5.0 / 3.0;              // eos.cpp:37-39
  cfg->abar = 1.0;
  cfg->cfl = 0.8;                      // hydro.cpp:90
  cfg->nspecies = 1;                   // material.cpp:34 default "single"
  cfg->mode_init = KM_EOS_INIT_DENS_PRES;  // eos.cpp:50 default "dens_pres"
  std::strncpy(cfg->species0, "single", sizeof(cfg->species0) - 1);
  // parthenon/time (kamayan_driver.cpp:28-53)
  cfg->dt_factor = 2.0;
  cfg->dt_init = DBL_MAX;
  cfg->dt_init_force = 1;
  cfg->dt_ceil = DBL_MAX;
  cfg->dt_floor = DBL_MIN;
  cfg->dt_min = -DBL_MAX;
  cfg->dt_min_cycle_limit = 10;
  cfg->dt_max = DBL_MAX;
  cfg->dt_max_cycle_limit = 1;
  cfg->pack_size = -1;
  cfg->dt_user = DBL_MAX;
  cfg->dt_force = -DBL_MAX;
  return KM_OK;
}

int km_cfg_set(km_cfg_t *cfg, const char *block_key, const char *value) {
  if (!cfg || !block_key || !value) return KM_EINVAL;
  const std::string key = trim(block_key), val = trim(value);
  if (key == "hydro/reconstruction") {
    static const char *const n[] = {"fog", "plm", "ppm", "wenoz"};
    static const int v[] = {KM_RECON_FOG, KM_RECON_PLM, KM_RECON_PPM, KM_RECON_WENOZ};
    return pick(key, val, n, v, &cfg->recon);
  }
  if (key == "hydro/slope_limiter") {
    static const char *const n[] = {"minmod", "van_leer", "mc"};
    static const int v[] = {KM_LIM_MINMOD, KM_LIM_VAN_LEER, KM_LIM_MC};
    return pick(key, val, n, v, &cfg->limiter);
  }
  if (key == "hydro/riemann") {  // hydro.cpp:66-69 registers hll and hllc only
    static const char *const n[] = {"hll", "hllc"};
    static const int v[] = {KM_RIEMANN_HLL, KM_RIEMANN_HLLC};
    return pick(key, val, n, v, &cfg->riemann);
  }
  if (key == "hydro/ReconstructionVars") {
    static const char *const n[] = {"primitive"};
    static const int v[] = {KM_RECONVARS_PRIMITIVE};
    return pick(key, val, n, v, &cfg->recon_vars);
  }
  if (key == "hydro/ReconstructionStrategy") {
    static const char *const n[] = {"scratchpad", "scratchvar"};
    static const int v[] = {KM_STRATEGY_SCRATCHPAD, KM_STRATEGY_SCRATCHVAR};
    return pick(key, val, n, v, &cfg->strategy);
  }
  if (key == "hydro/EMF_averaging") {
    static const char *const n[] = {"arithmetic"};
    static const int v[] = {KM_EMF_ARITHMETIC};
    return pick(key, val, n, v, &cfg->emf_averaging);
  }
  if (key == "hydro/cfl") {
    if (!parse_real(val, &cfg->cfl)) {
      km::set_error("invalid value '" + val + "' for " + key);
      return KM_EINVAL;
    }
    return KM_OK;
  }
  if (key == "physics/fluid") {  // physics.cpp:20-21
    static const char *const n[] = {"1t", "3t"};
    static const int v[] = {KM_FLUID_ONET, KM_FLUID_THREET};
    return pick(key, val, n, v, &cfg->fluid);
  }
  if (key == "physics/MHD") {
    static const char *const n[] = {"off", "ct"};
    static const int v[] = {KM_MHD_OFF, KM_MHD_CT};
    return pick(key, val, n, v, &cfg->mhd);
  }
  if (key == "geometry/geometry") {
    static const char *const n[] = {"cartesian", "cylindrical"};
    static const int v[] = {KM_GEOM_CARTESIAN, KM_GEOM_CYLINDRICAL};
    return pick(key, val, n, v, &cfg->geometry);
  }
  if (key == "parthenon/time/integrator") {
    static const char *const n[] = {"rk1", "rk2", "rk3"};
    static const int v[] = {1, 2, 3};
    return pick(key, val, n, v, &cfg->nstages);
  }
  if (key == "eos/mode_init") {  // eos.cpp:46-62
    static const char *const n[] = {"dens_pres", "dens_ener", "dens_temp"};
    static const int v[] = {KM_EOS_INIT_DENS_PRES, KM_EOS_INIT_DENS_ENER, KM_EOS_INIT_DENS_TEMP};
    return pick(key, val, n, v, &cfg->mode_init);
  }
  // material/<species>/gamma and Abar (eos.cpp:32-40, material.cpp:55-60): the EOS is built
  // from the first species (MakeEos, eos.cpp:170-181); other species' values are kept by
  // the reference but unused, so they are accepted and ignored here
  if (key.compare(0, 9, "material/") == 0 &&
      (ends_with(key, "/gamma") || ends_with(key, "/Abar") || ends_with(key, "/Z") ||
       ends_with(key, "/eos_type"))) {
    const std::string spec = key.substr(9, key.rfind('/') - 9);
    const std::string sub = key.substr(key.rfind('/') + 1);
    if (sub == "eos_type") {
      if (val != "gamma") {
        km::set_error("invalid value '" + val + "' for " + key);
        return KM_EINVAL;
      }
      return KM_OK;
    }
    double x = 0.;
    if (!parse_real(val, &x)) {
      km::set_error("invalid value '" + val + "' for " + key);
      return KM_EINVAL;
    }
    if (spec == cfg->species0) {
      if (sub == "gamma") cfg->gamma = x;
      if (sub == "Abar") cfg->abar = x;
    }
    return KM_OK;
  }
  {
    // parthenon/time step control (kamayan_driver.cpp:28-53)
    struct RealKey {
      const char *name;
      double *dst;
    } reals[] = {{"parthenon/time/dt_factor", &cfg->dt_factor},
                 {"parthenon/time/dt_init", &cfg->dt_init},
                 {"parthenon/time/dt_ceil", &cfg->dt_ceil},
                 {"parthenon/time/dt_floor", &cfg->dt_floor},
                 {"parthenon/time/dt_min", &cfg->dt_min},
                 {"parthenon/time/dt_max", &cfg->dt_max},
                 {"parthenon/time/dt_user", &cfg->dt_user},
                 {"parthenon/time/dt_force", &cfg->dt_force}};
    for (const auto &r : reals)
      if (key == r.name) {
        if (!parse_real(val, r.dst)) {
          km::set_error("invalid value '" + val + "' for " + key);
          return KM_EINVAL;
        }
        return KM_OK;
      }
    struct IntKey {
      const char *name;
      int32_t *dst;
    } ints[] = {{"parthenon/time/dt_min_cycle_limit", &cfg->dt_min_cycle_limit},
                {"parthenon/time/dt_max_cycle_limit", &cfg->dt_max_cycle_limit},
                {"parthenon/mesh/pack_size", &cfg->pack_size}};
    for (const auto &r : ints)
      if (key == r.name) {
        long iv = 0;
        // pack_size: -1 (every block in one pack) or a positive block count
        const bool pack = r.dst == &cfg->pack_size;
        if (!parse_int(val, &iv) || iv < (pack ? -1 : 0) || (pack && iv == 0) ||
            iv > 1000000000L) {
          km::set_error("invalid value '" + val + "' for " + key);
          return KM_EINVAL;
        }
        *r.dst = (int32_t)iv;
        return KM_OK;
      }
    if (key == "parthenon/time/dt_init_force") {  // the deck writer prints True/False
      std::string low = val;
      for (auto &ch : low) ch = (char)std::tolower(static_cast<unsigned char>(ch));
      if (low == "true" || low == "1") cfg->dt_init_force = 1;
      else if (low == "false" || low == "0") cfg->dt_init_force = 0;
      else {
        km::set_error("invalid value '" + val + "' for " + key);
        return KM_EINVAL;
      }
      return KM_OK;
    }
  }
  if (key == "material/species") {  // comma-separated, unique (material.cpp:34-43)
    int n = 0;
    std::string item;
    std::istringstream in(val);
    std::vector<std::string> seen;
    while (std::getline(in, item, ',')) {
      item = trim(item);
      for (const auto &x : seen)
        if (x == item) {
          km::set_error("material/species must be unique");
          return KM_EINVAL;
        }
      seen.push_back(item);
      ++n;
    }
    if (n < 1 || n > 64 || seen[0].empty() || seen[0].size() >= sizeof(cfg->species0)) {
      km::set_error("invalid value '" + val + "' for " + key);
      return KM_EINVAL;
    }
    cfg->nspecies = n;
    std::memset(cfg->species0, 0, sizeof(cfg->species0));
    std::memcpy(cfg->species0, seen[0].data(), seen[0].size());
    return KM_OK;
  }
  if (key == "kamayan_amd/precision") {
    static const char *const n[] = {"fp64", "fp32_mixed", "fp64_fma"};
    static const int v[] = {KM_PREC_FP64, KM_PREC_FP32_MIXED, KM_PREC_FP64_FMA};
    return pick(key, val, n, v, &cfg->precision);
  }
  // ndim follows from parthenon/mesh/nx2 and nx3 together; km_cfg_parse_deck and
  // km_sim_create (from km_mesh_desc_t) set it.
  return KM_OK;  // keys outside this path are kept by the caller, not validated here
}

int km_cfg_parse_deck(km_cfg_t *cfg, const char *deck_text) {
  if (!cfg || !deck_text) return KM_EINVAL;
  std::istringstream in(deck_text);
  std::string line, block;
  long nx2 = -1, nx3 = -1;
  std::vector<std::pair<std::string, std::string>> kv;
  while (std::getline(in, line)) {
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line = line.substr(0, hash);
    line = trim(line);
    if (line.empty()) continue;
    if (line.front() == '<') {
      const size_t close = line.find('>');
      if (close == std::string::npos) {
        km::set_error("malformed block header: " + line);
        return KM_EINVAL;
      }
      block = trim(line.substr(1, close - 1));
      continue;
    }
    const size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    const std::string key = block + "/" + trim(line.substr(0, eq));
    const std::string val = trim(line.substr(eq + 1));
    if (key == "parthenon/mesh/nx2") {
      parse_int(val, &nx2);
      continue;
    }
    if (key == "parthenon/mesh/nx3") {
      parse_int(val, &nx3);
      continue;
    }
    kv.emplace_back(key, val);
  }
  // material/species first: it names the species whose gamma and Abar the EOS takes,
  // wherever the deck lists it
  for (int pass = 0; pass < 2; ++pass)
    for (const auto &e : kv) {
      if ((e.first == "material/species") != (pass == 0)) continue;
      const int rc = km_cfg_set(cfg, e.first.c_str(), e.second.c_str());
      if (rc != KM_OK) return rc;
    }
  if (nx2 >= 1 || nx3 >= 1) {
    const long n2 = nx2 >= 1 ? nx2 : 32, n3 = nx3 >= 1 ? nx3 : 32;  // grid.cpp:51-55
    cfg->ndim = n3 > 1 ? 3 : (n2 > 1 ? 2 : 1);
  }
  return KM_OK;
}

// Adaptive-refinement keys: parthenon/mesh/{numlevel,derefine_count} and the
// <kamayan/refinementN> blocks of grid::SetupParams (grid.cpp:96-124); a criterion's
// max_level defaults to numlevel, also when numlevel is set after the field.
int km_amr_adapt_default(km_amr_adapt_t *a) {
  if (!a) return KM_EINVAL;
  std::memset(a, 0, sizeof(*a));
  a->numlevel = 1;         // grid.cpp:48
  a->derefine_count = 10;  // Parthenon's parthenon/mesh/derefine_count default
  return KM_OK;
}

int km_amr_adapt_set(km_amr_adapt_t *a, int32_t mhd, const char *block_key, const char *value) {
  if (!a || !block_key || !value) return KM_EINVAL;
  const std::string key = trim(block_key), val = trim(value);
  long iv = 0;
  double rv = 0.;
  if (key == "parthenon/mesh/numlevel" || key == "parthenon/mesh/derefine_count") {
    if (!parse_int(val, &iv) || iv < (key.back() == 'l' ? 1 : 0) || iv > 1000) {
      km::set_error("invalid value '" + val + "' for " + key);
      return KM_EINVAL;
    }
    if (key.back() == 'l') {
      for (int c = 0; c < a->ncrit; ++c)
        if (a->crit[c].max_level == a->numlevel) a->crit[c].max_level = (int32_t)iv;
      a->numlevel = (int32_t)iv;
    } else {
      a->derefine_count = (int32_t)iv;
    }
    return KM_OK;
  }
  const std::string pre = "kamayan/refinement";
  if (key.compare(0, pre.size(), pre) != 0) return KM_OK;  // unrelated key
  const size_t slash = key.find('/', pre.size());
  long n = -1;
  if (slash == std::string::npos || !parse_int(key.substr(pre.size(), slash - pre.size()), &n) ||
      n < 0 || n >= KM_AMR_MAX_CRIT) {
    km::set_error("unsupported refinement block in " + key);
    return KM_EINVAL;
  }
  const std::string sub = key.substr(slash + 1);
  km_amr_criterion_t &c = a->crit[n];
  if (n >= a->ncrit) {
    if (sub != "field") return KM_OK;  // the criterion starts with its field
    for (int k = a->ncrit; k <= n; ++k) {
      a->crit[k].var = -1;
      a->crit[k].max_level = a->numlevel;  // grid.cpp:112-113
      a->crit[k].refine_tol = 0.8;         // grid.cpp:107
      a->crit[k].derefine_tol = 0.2;       // grid.cpp:108-109
      a->crit[k].filter = 0.01;            // grid.cpp:110-111
    }
    a->ncrit = (int32_t)n + 1;
  }
  if (sub == "field") {
    const int nc = mhd == KM_MHD_CT ? 8 : 5;
    int base = -1;
    if (val == "dens") base = KM_DENS;
    else if (val == "momentum") base = KM_MOM0;
    else if (val == "ener") base = KM_ENER;
    else if (val == "magc" && mhd == KM_MHD_CT) base = KM_MAGC0;
    else if (val == "velocity") base = nc + 0;
    else if (val == "pres") base = nc + 3;
    else if (val == "bulk modulus") base = nc + 4;
    else if (val == "eint") base = nc + 5;
    else if (val == "temp") base = nc + 6;
    if (base < 0) {
      km::set_error("unknown refinement field '" + val + "'");
      return KM_EINVAL;
    }
    c.var = base;
    return KM_OK;
  }
  if (sub == "comp4") {
    if (!parse_int(val, &iv) || iv < 0 || iv > 2 || c.var < 0) {
      km::set_error("invalid value '" + val + "' for " + key);
      return KM_EINVAL;
    }
    c.var += (int32_t)iv;
    return KM_OK;
  }
  if (sub == "method") {
    if (val != "loehner") {
      km::set_error("refinement method '" + val +
                    "' is not provided (Parthenon's built-in criteria are not part of this "
                    "library)");
      return KM_EINVAL;
    }
    return KM_OK;
  }
  if (sub == "max_level") {
    if (!parse_int(val, &iv) || iv < 0) {
      km::set_error("invalid value '" + val + "' for " + key);
      return KM_EINVAL;
    }
    c.max_level = (int32_t)iv;
    return KM_OK;
  }
  if (sub == "refine_tol" || sub == "derefine_tol" || sub == "filter") {
    if (!parse_real(val, &rv)) {
      km::set_error("invalid value '" + val + "' for " + key);
      return KM_EINVAL;
    }
    (sub == "refine_tol" ? c.refine_tol : sub == "derefine_tol" ? c.derefine_tol : c.filter) = rv;
    return KM_OK;
  }
  return KM_OK;
}

int km_cfg_min_ghost(const km_cfg_t *cfg) {
  if (!cfg) return -1;
  static const int ng[] = {1, 2, 3, 3};  // Hydro.py:14 {fog, plm, ppm, wenoz}
  int g = ng[cfg->recon < 0 || cfg->recon > 3 ? 0 : cfg->recon];
  if (cfg->mhd == KM_MHD_CT) g += 1;  // widened flux bounds, hydro_add_flux_tasks.cpp:59-70
  return g;
}

}  // extern "C"
