// Calibration of recorded values to volts or physical units
// (core/calibrate.nim): the rules' index (initCalibrationPlan) and each
// row's factor (resolveRow), on the host.  Both are per-row metadata work; the
// samples are scaled on the device, by a step of the row kernel
// (pseg_plan_apply_calib) with the multipliers pseg_plan_calib_factors
// derives from these factors.
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "pseg_plan.hpp"

namespace pseg {

namespace {

// The criteria bits of a rule, ranked by specificity (core/calibrate.nim:
// 217-224): station > channel > channel set > component > sensor type.
constexpr uint8_t bSensorType = 1, bComponent = 2, bChannelSet = 4, bChannel = 8, bStation = 16;
constexpr double kConflictTolerance = 1e-3;

// hundredths (core/calibrate.nim:225): Nim's round is half away from zero,
// as std::round; + 0.0 folds -0 into 0 (the reference hashes -0.0 as 0.0).
double hundredths(double x) { return std::round(x * 100) / 100 + 0.0; }

// A rule's criteria, or a row's values under one criteria set; fields
// outside the mask are zero so that map order and equality ignore them
// (MatchKey's hash and ==, core/calibrate.nim:227-248).
struct MatchKey {
  uint8_t mask = 0;
  int sensor_type = 0, component = 0, channel_set = 0;
  double channel = 0, line = 0, point = 0;
  int index = 0;
  bool operator<(const MatchKey &o) const {
    return std::tie(mask, sensor_type, component, channel_set, channel, line, point, index) <
           std::tie(o.mask, o.sensor_type, o.component, o.channel_set, o.channel, o.line, o.point,
                    o.index);
  }
};

bool finite(double x) { return std::isfinite(x); }

std::string num(double x) {
  // Nim's $ of a float64: "20.0", "1e-06", "0.5"
  char b[64];
  if (std::isnan(x)) return "nan";
  if (std::isinf(x)) return x > 0 ? "inf" : "-inf";
  if (x == std::floor(x) && std::fabs(x) < 1e16) {
    snprintf(b, sizeof b, "%.1f", x);
  } else {
    snprintf(b, sizeof b, "%.16g", x);
    // the shortest text that reads back the same value
    for (int p = 1; p <= 17; p++) {
      char t[64];
      snprintf(t, sizeof t, "%.*g", p, x);
      if (std::strtod(t, nullptr) == x) {
        snprintf(b, sizeof b, "%s", t);
        break;
      }
    }
  }
  return b;
}

std::string fmt(const char *f, ...) {
  char b[512];
  va_list a;
  va_start(a, f);
  vsnprintf(b, sizeof b, f, a);
  va_end(a);
  return b;
}

// `$` of a RuleMatch (core/calibrate.nim:266-276).
std::string match_text(const pseg_calib_rule &r) {
  std::string out;
  auto add = [&](const std::string &p) { out += (out.empty() ? "" : ", ") + p; };
  if (r.has_station)
    add("station " + num(r.station_line) + "/" + num(r.station_point) +
        (r.station_index != 0 ? "." + std::to_string(r.station_index) : ""));
  if (!std::isnan(r.channel)) add("channel " + num(r.channel));
  if (r.channel_set >= 0) add("channel set " + std::to_string(r.channel_set));
  if (r.component >= 0) add("component " + std::to_string(r.component));
  if (r.sensor_type >= 0) add("sensor type " + std::to_string(r.sensor_type));
  return out.empty() ? "any row" : out;
}

bool is_quantity(int u) {
  return u == PSEG_AU_VELOCITY || u == PSEG_AU_ACCELERATION || u == PSEG_AU_PRESSURE ||
         u == PSEG_AU_FORCE;
}

// expectedUnit (core/calibrate.nim:412-417): the quantity a SEG-D sensor
// type measures.
int expected_unit(uint8_t sensor_type) {
  if (sensor_type == 1) return PSEG_AU_PRESSURE;
  if (sensor_type >= 2 && sensor_type <= 5) return PSEG_AU_VELOCITY;
  if (sensor_type >= 6 && sensor_type <= 9) return PSEG_AU_ACCELERATION;
  return PSEG_AU_RECORDED;
}

bool differs(double a, double b) {
  return std::fabs(a - b) > kConflictTolerance * std::max(std::fabs(a), std::fabs(b));
}

// usable (core/calibrate.nim:422-428): the factor and its inverse fit a
// float32.
bool usable(double f) {
  const double top = (double)FLT_MAX;
  return f > 0 && f <= top && 1 / f <= top;
}

}  // namespace

struct CalibPlan {
  int target = PSEG_CAL_PHYSICAL;
  int file_factors = PSEG_FF_FALLBACK;
  bool scale_aux = true;
  std::vector<uint8_t> classes;  // criteria sets of the specific rules, most specific first
  int catch_all = -1;
  std::map<MatchKey, int> index;
  std::vector<double> descale, sens, gain;  // per rule; NaN where it gives none (gain linear)
  std::vector<int> sens_unit;
};

std::shared_ptr<CalibPlan> calib_new() { return std::make_shared<CalibPlan>(); }

std::string calib_compile(const pseg_calib_rule *rules, int n, int target, int file_factors,
                          int scale_aux, CalibPlan &P) {
  if (target != PSEG_CAL_PHYSICAL && target != PSEG_CAL_VOLTS)
    return fmt("unknown calibration target %d", target);
  if (file_factors < PSEG_FF_FALLBACK || file_factors > PSEG_FF_IGNORE)
    return fmt("unknown file factor order %d", file_factors);
  if (n < 0 || (n > 0 && !rules)) return "NULL rules";
  P = CalibPlan();
  P.target = target;
  P.file_factors = file_factors;
  P.scale_aux = scale_aux != 0;
  P.descale.assign(n, NAN);
  P.sens.assign(n, NAN);
  P.gain.assign(n, NAN);
  P.sens_unit.assign(n, PSEG_AU_RECORDED);
  bool masks[32] = {false};
  for (int i = 0; i < n; i++) {
    const pseg_calib_rule &r = rules[i];
    auto bad = [&](const std::string &m) { return fmt("rule %d: %s", i, m.c_str()); };
    // descale() and sensitivity() refuse these when a rule is built
    // (core/units.nim:185-207, 331-341); here the factors arrive in SI
    const bool has_d = !std::isnan(r.descale), has_s = !std::isnan(r.sensitivity);
    if (has_d && !(finite(r.descale) && r.descale > 0))
      return bad("a descale must be finite and positive, got " + num(r.descale));
    if (has_s && !(finite(r.sensitivity) && r.sensitivity > 0))
      return bad("a sensitivity must be finite and positive, got " + num(r.sensitivity));
    if (has_s && !is_quantity(r.sensitivity_unit))
      return bad(fmt("a sensitivity needs the quantity it measures (unit %d)", r.sensitivity_unit));
    // initCalibrationPlan's checks (core/calibrate.nim:328-352)
    if (!(has_d || has_s || !std::isnan(r.gain_db)))
      return bad("gives no factor (set a descale, a sensitivity or gainDb)");
    if (!std::isnan(r.gain_db) && !finite(r.gain_db))
      return bad("gainDb must be finite, got " + num(r.gain_db));
    double lin = NAN;
    if (!std::isnan(r.gain_db)) {
      lin = std::pow(10.0, r.gain_db / 20);
      if (!(finite(lin) && lin > 0))
        return bad("gainDb = " + num(r.gain_db) + " is out of range: 10^(gainDb/20) = " + num(lin));
    }
    if (r.sensor_type > 255 || r.sensor_type < -1)
      return bad(fmt("sensorType must be -1 (any) or 0 .. 255, got %d", r.sensor_type));
    if (r.component > 255 || r.component < -1)
      return bad(fmt("component must be -1 (any) or 0 .. 255, got %d", r.component));
    if (r.channel_set < -1)
      return bad(fmt("channelSet must be -1 (any) or ≥ 0, got %d", r.channel_set));
    if (!std::isnan(r.channel) && !finite(r.channel))
      return bad("channel must be finite (NaN: any), got " + num(r.channel));
    if (r.has_station && !(finite(r.station_line) && finite(r.station_point)))
      return bad("the station's line and point must be finite");
    // keyOf (core/calibrate.nim:256-264)
    MatchKey key;
    if (r.sensor_type >= 0) key.mask |= bSensorType, key.sensor_type = r.sensor_type;
    if (r.component >= 0) key.mask |= bComponent, key.component = r.component;
    if (r.channel_set >= 0) key.mask |= bChannelSet, key.channel_set = r.channel_set;
    if (!std::isnan(r.channel)) key.mask |= bChannel, key.channel = hundredths(r.channel);
    if (r.has_station) {
      key.mask |= bStation;
      key.line = hundredths(r.station_line);
      key.point = hundredths(r.station_point);
      key.index = (int16_t)r.station_index;
    }
    if (key.mask == 0) {
      if (P.catch_all >= 0)
        return bad(fmt("has no criteria, like rule %d: one catch-all only", P.catch_all));
      P.catch_all = i;
    } else {
      auto it = P.index.find(key);
      if (it != P.index.end())
        return bad(fmt("matches the same rows as rule %d (%s)", it->second, match_text(r).c_str()));
      P.index[key] = i;
      masks[key.mask] = true;
    }
    P.descale[i] = has_d ? r.descale : NAN;
    P.sens[i] = has_s ? r.sensitivity : NAN;
    P.sens_unit[i] = has_s ? r.sensitivity_unit : PSEG_AU_RECORDED;
    P.gain[i] = lin;
  }
  for (int m = 31; m >= 1; m--)
    if (masks[m]) P.classes.push_back((uint8_t)m);
  return "";
}

namespace {

struct Found {
  int d = -1, s = -1, g = -1;
};

void take(const CalibPlan &p, int r, Found &f) {
  if (f.d < 0 && !std::isnan(p.descale[r])) f.d = r;
  if (f.s < 0 && !std::isnan(p.sens[r])) f.s = r;
  if (f.g < 0 && !std::isnan(p.gain[r])) f.g = r;
}

int lookup(const CalibPlan &p, const MatchKey &k) {
  auto it = p.index.find(k);
  return it == p.index.end() ? -1 : it->second;
}

// specificRules (core/calibrate.nim:383-410): the most specific rule giving
// each factor, the catch-all left out.
Found specific_rules(const CalibPlan &p, const pseg_calib_row &k) {
  Found f;
  for (uint8_t mask : p.classes) {
    MatchKey key;
    key.mask = mask;
    if (mask & bSensorType) key.sensor_type = k.sensor_type;
    if (mask & bComponent) key.component = k.component;
    if (mask & bChannelSet) {
      if (k.channel_set < 0) continue;
      key.channel_set = k.channel_set;
    }
    if (mask & bChannel) {
      if (std::isnan(k.channel)) continue;
      key.channel = hundredths(k.channel);
    }
    if (mask & bStation) {
      if (std::isnan(k.rcv_line) || std::isnan(k.rcv_point)) continue;
      key.line = hundredths(k.rcv_line);
      key.point = hundredths(k.rcv_point);
      // the exact index, then a rule for any index (0); a trace without an
      // index (0) also matches index 1
      const int16_t idx = (int16_t)k.rcv_index;
      const int16_t tries[2] = {idx != 0 ? idx : (int16_t)0, idx != 0 ? (int16_t)0 : (int16_t)1};
      for (int16_t t : tries) {
        key.index = t;
        const int r = lookup(p, key);
        if (r >= 0) take(p, r, f);
      }
    } else {
      const int r = lookup(p, key);
      if (r >= 0) take(p, r, f);
    }
    if (f.d >= 0 && f.s >= 0 && f.g >= 0) break;
  }
  return f;
}

}  // namespace

// resolveRow (core/calibrate.nim:430-497).
pseg_row_calibration calib_resolve_row(const CalibPlan &p, const pseg_calib_row &k,
                                       bool has_file_scales) {
  pseg_row_calibration R;
  R.factor = 1;
  R.unit = PSEG_AU_RECORDED;
  R.status = PSEG_CS_KEPT;
  R.descale_from = R.sensitivity_from = PSEG_FS_NONE;
  R.descale_rule = R.sensitivity_rule = R.gain_rule = -1;
  R.conflict = R.mismatch = 0;
  if (k.failed) return R;
  const bool aux = k.status == 3;
  const bool live = k.status == 0 || k.status == 1;
  if (aux && !p.scale_aux) return R;
  const bool use_file = p.file_factors != PSEG_FF_IGNORE && has_file_scales;
  // initFileScale() where the file is not used
  const double fs_descale = use_file ? k.file_descale : NAN;
  const double fs_sens = use_file ? k.file_sensitivity : NAN;
  const int fs_unit = use_file ? k.file_unit : PSEG_AU_RECORDED;
  const bool fs_physical = use_file && k.file_physical;
  const Found spec = specific_rules(p, k);
  const int c = p.catch_all;
  const int rd = spec.d >= 0 ? spec.d : (c >= 0 && !std::isnan(p.descale[c]) ? c : -1);
  const int rs = spec.s >= 0 ? spec.s : (c >= 0 && !std::isnan(p.sens[c]) ? c : -1);
  const int rg = spec.g >= 0 ? spec.g : (c >= 0 && !std::isnan(p.gain[c]) ? c : -1);
  int unit = PSEG_AU_RECORDED;
  double factor = NAN;
  if (use_file && fs_physical) {
    // the headers say the samples are already physical
    R.conflict = rd >= 0 || rs >= 0 || rg >= 0;
    if (p.target == PSEG_CAL_PHYSICAL && !aux && finite(fs_descale) && fs_descale > 0) {
      factor = fs_descale;
      unit = fs_unit;
      R.descale_from = R.sensitivity_from = PSEG_FS_FILE;
    }
  } else {
    double d = NAN, s = NAN;
    int s_unit = PSEG_AU_RECORDED;
    const bool file_d = finite(fs_descale) && fs_descale > 0;
    const bool file_s = finite(fs_sens) && fs_sens > 0 && is_quantity(fs_unit);
    // each factor in the order fileFactors sets
    auto pick = [&](bool file_has, double file_v, int spec_idx, int rule_idx,
                    const std::vector<double> &field, double &v, int32_t &src, int32_t &ri) {
      if (p.file_factors == PSEG_FF_FIRST && file_has) {
        v = file_v;
        src = PSEG_FS_FILE;
      } else if (spec_idx >= 0) {
        v = field[spec_idx];
        src = PSEG_FS_RULE;
        ri = spec_idx;
      } else if (file_has) {
        v = file_v;
        src = PSEG_FS_FILE;
      } else if (rule_idx >= 0) {
        v = field[rule_idx];
        src = PSEG_FS_RULE;
        ri = rule_idx;
      }
    };
    pick(file_d, fs_descale, spec.d, rd, p.descale, d, R.descale_from, R.descale_rule);
    const bool want_s = !aux && p.target == PSEG_CAL_PHYSICAL;
    if (want_s) {
      pick(file_s, fs_sens, spec.s, rs, p.sens, s, R.sensitivity_from, R.sensitivity_rule);
      s_unit = R.sensitivity_from == PSEG_FS_FILE ? fs_unit
               : R.sensitivity_rule >= 0          ? p.sens_unit[R.sensitivity_rule]
                                                  : PSEG_AU_RECORDED;
    }
    // a conflict: the file and a rule both give a factor, and they differ
    if (file_d && rd >= 0 && differs(fs_descale, p.descale[rd])) R.conflict = 1;
    if (want_s && file_s && rs >= 0 && (fs_unit != p.sens_unit[rs] || differs(fs_sens, p.sens[rs])))
      R.conflict = 1;
    const double g = rg >= 0 ? p.gain[rg] : 1.0;
    R.gain_rule = rg;
    if (want_s && !std::isnan(s)) {
      const int want = expected_unit(k.sensor_type);
      R.mismatch = (want != PSEG_AU_RECORDED && want != s_unit) ||
                   (is_quantity(fs_unit) && fs_unit != s_unit);
    }
    if (!std::isnan(d)) {
      if (aux || p.target == PSEG_CAL_VOLTS) {
        factor = d / g;
        unit = PSEG_AU_VOLT;
      } else if (!std::isnan(s)) {
        factor = d / (g * s);
        unit = s_unit;
      }
    }
  }
  if (usable(factor)) {
    R.factor = factor;
    R.unit = unit;
    R.status = PSEG_CS_CONVERTED;
  } else if (live) {
    R.factor = NAN;
    R.status = PSEG_CS_UNRESOLVED;
  }
  return R;
}

// count (core/calibrate.nim:499-512).
void calib_count(pseg_calib_summary &s, const pseg_row_calibration &r, const CalibPlan &p) {
  switch (r.status) {
    case PSEG_CS_CONVERTED:
      s.converted++;
      s.units[r.unit]++;
      break;
    case PSEG_CS_UNRESOLVED:
      s.unresolved++;
      if (r.descale_from == PSEG_FS_NONE) s.no_descale++;
      else if (p.target == PSEG_CAL_PHYSICAL && r.sensitivity_from == PSEG_FS_NONE)
        s.no_sensitivity++;
      break;
    default:
      s.kept++;
      s.units[PSEG_AU_RECORDED]++;
      break;
  }
  if (r.conflict) s.conflicts++;
  if (r.mismatch) s.mismatches++;
}

}  // namespace pseg

struct pseg_calib {
  pseg::CalibPlan plan;
};

using namespace pseg;

pseg_status pseg_calib_compile(const pseg_calib_rule *rules, int32_t n_rules, int32_t target,
                               int32_t file_factors, int32_t scale_auxiliary, pseg_calib **out) {
  if (!out) return set_error(PSEG_E_INVALID, "out is NULL");
  *out = nullptr;
  pseg_calib *c = new pseg_calib;
  const std::string e = calib_compile(rules, n_rules, target, file_factors, scale_auxiliary, c->plan);
  if (!e.empty()) {
    delete c;
    return set_error(PSEG_E_INVALID, e);
  }
  *out = c;
  return PSEG_OK;
}

pseg_status pseg_calib_destroy(pseg_calib *cal) {
  delete cal;
  return PSEG_OK;
}

pseg_status pseg_calib_resolve(const pseg_calib *cal, const pseg_calib_row *rows, int32_t n_rows,
                               int32_t has_file_scales, pseg_row_calibration *out,
                               pseg_calib_summary *summary) {
  if (!cal) return set_error(PSEG_E_INVALID, "a calibration must come from pseg_calib_compile");
  if (n_rows < 0) return set_error(PSEG_E_INVALID, "n_rows must be >= 0");
  if (n_rows > 0 && (!rows || !out)) return set_error(PSEG_E_INVALID, "NULL rows or out");
  pseg_calib_summary s{};
  for (int32_t r = 0; r < n_rows; r++) {
    out[r] = calib_resolve_row(cal->plan, rows[r], has_file_scales != 0);
    calib_count(s, out[r], cal->plan);
  }
  if (summary) *summary = s;
  return PSEG_OK;
}

pseg_status pseg_plan_calib_factors(const pseg_plan *P, const pseg_calib_row *rows,
                                    int32_t n_rows, int32_t has_file_scales, double *factors,
                                    pseg_calib_summary *summaries, int32_t *n_calib) {
  if (!P) return set_error(PSEG_E_INVALID, "a plan must come from pseg_plan_compile");
  if (n_calib) *n_calib = P->n_calib;
  if (!factors && !summaries) return PSEG_OK;
  if (n_rows < 0) return set_error(PSEG_E_INVALID, "n_rows must be >= 0");
  if (n_rows > 0 && !rows) return set_error(PSEG_E_INVALID, "NULL rows");
  for (int c = 0; c < P->n_calib; c++) {
    const CalibPlan &cp = *P->calib[c];
    pseg_calib_summary s{};
    for (int32_t r = 0; r < n_rows; r++) {
      const pseg_row_calibration rc = calib_resolve_row(cp, rows[r], has_file_scales != 0);
      calib_count(s, rc, cp);
      // calibrateImpl's k (core/calibrate.nim:615-624): NaN stays NaN, a
      // kept row's 1 stays 1
      const double k = rc.status == PSEG_CS_CONVERTED && P->calib_inverse[c] ? 1 / rc.factor
                                                                               : rc.factor;
      if (factors) factors[(size_t)c * n_rows + r] = k;
    }
    if (summaries) summaries[c] = s;
  }
  return PSEG_OK;
}
