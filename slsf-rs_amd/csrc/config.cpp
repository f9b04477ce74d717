// config.cpp -- job configuration: a reader for the TOML subset the reference's
// job files use, the reference's schema with its aliases and defaults
// (src/model.rs:159-213,518-644), task expansion with the reference's names,
// seeds and disorder draws (src/model.rs:697-828), durations
// (src/model.rs:1033-1060) and round-robin rank selection (src/model.rs:62-69).
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/slsf.h"
#include "common.h"

using namespace slsf;

namespace {

// ---------------------------------------------------------------------------
// SplitMix64 FastRng (src/types.rs:9-82) for the host-side disorder draws
// ---------------------------------------------------------------------------
struct SplitMix64 {
  uint64_t state;
  explicit SplitMix64(uint64_t seed) : state(seed) {}
  uint64_t next_u64() {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  bool next_bool() { return (int32_t)(uint32_t)next_u64() < 0; }       // rand bool
  double next_f64() { return (double)(next_u64() >> 11) * 0x1.0p-53; }  // rand f64
};

// ---------------------------------------------------------------------------
// TOML subset: tables, bare keys, strings, integers, floats, booleans, arrays
// ---------------------------------------------------------------------------
struct Value {
  enum Kind { STR, INT, FLT, BOOL, ARR } kind = INT;
  std::string s;
  int64_t i = 0;
  double f = 0.0;
  bool b = false;
  std::vector<Value> a;
};

using Table = std::map<std::string, Value>;

struct Parser {
  const char* p;
  int line = 1;
  std::string err;

  bool error(const std::string& m) {
    if (err.empty()) err = "TOML parse error at line " + std::to_string(line) + ": " + m;
    return false;
  }
  void skip_ws() {
    while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
  }
  void skip_comment() {
    if (*p == '#')
      while (*p && *p != '\n') ++p;
  }
  void skip_ws_nl_comments() {
    for (;;) {
      skip_ws();
      skip_comment();
      if (*p == '\n') { ++line; ++p; continue; }
      break;
    }
  }
  bool parse_key(std::string& k) {
    skip_ws();
    if (*p == '"') return parse_string(k);
    const char* s = p;
    while (*p && (isalnum((unsigned char)*p) || *p == '_' || *p == '-')) ++p;
    if (p == s) return error("expected a key");
    k.assign(s, p);
    return true;
  }
  bool parse_string(std::string& out) {
    ++p;  // opening quote
    out.clear();
    while (*p && *p != '"') {
      if (*p == '\n') return error("unterminated string");
      if (*p == '\\') {
        ++p;
        switch (*p) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case '"': out += '"'; break;
          case '\\': out += '\\'; break;
          default: return error("unsupported escape in string");
        }
        ++p;
      } else {
        out += *p++;
      }
    }
    if (*p != '"') return error("unterminated string");
    ++p;
    return true;
  }
  bool parse_value(Value& v) {
    skip_ws();
    if (*p == '"') {
      v.kind = Value::STR;
      return parse_string(v.s);
    }
    if (*p == '[') {
      ++p;
      v.kind = Value::ARR;
      for (;;) {
        skip_ws_nl_comments();
        if (*p == ']') { ++p; return true; }
        Value e;
        if (!parse_value(e)) return false;
        v.a.push_back(std::move(e));
        skip_ws_nl_comments();
        if (*p == ',') { ++p; continue; }
        if (*p == ']') { ++p; return true; }
        return error("expected ',' or ']' in array");
      }
    }
    if (!strncmp(p, "true", 4)) { p += 4; v.kind = Value::BOOL; v.b = true; return true; }
    if (!strncmp(p, "false", 5)) { p += 5; v.kind = Value::BOOL; v.b = false; return true; }
    // number: collect the token, drop '_' separators
    const char* s = p;
    while (*p && (isalnum((unsigned char)*p) || *p == '+' || *p == '-' || *p == '.' || *p == '_')) ++p;
    std::string tok;
    for (const char* q = s; q < p; ++q)
      if (*q != '_') tok += *q;
    if (tok.empty()) return error("expected a value");
    const std::string body = (tok[0] == '+' || tok[0] == '-') ? tok.substr(1) : tok;
    if (body == "inf" || body == "nan") {
      v.kind = Value::FLT;
      v.f = body == "inf" ? INFINITY : NAN;
      if (tok[0] == '-') v.f = -v.f;
      return true;
    }
    const bool is_float = tok.find_first_of(".eE") != std::string::npos;
    errno = 0;
    char* end = nullptr;
    if (is_float) {
      v.kind = Value::FLT;
      v.f = strtod(tok.c_str(), &end);
    } else {
      v.kind = Value::INT;
      v.i = strtoll(tok.c_str(), &end, 10);
    }
    if (!end || *end || errno) return error("invalid number '" + tok + "'");
    return true;
  }
  bool parse(std::map<std::string, Table>& doc) {
    std::string table;  // "" = root
    doc[table];
    for (;;) {
      skip_ws_nl_comments();
      if (!*p) return true;
      if (*p == '[') {
        ++p;
        std::string name;
        if (!parse_key(name)) return false;
        skip_ws();
        if (*p != ']') return error("expected ']' after table name");
        ++p;
        if (doc.count(name)) return error("duplicate table [" + name + "]");
        table = name;
        doc[table];
        continue;
      }
      std::string key;
      if (!parse_key(key)) return false;
      skip_ws();
      if (*p != '=') return error("expected '=' after key " + key);
      ++p;
      Value v;
      if (!parse_value(v)) return false;
      Table& t = doc[table];
      if (t.count(key)) return error("duplicate key " + key);
      t[key] = std::move(v);
      skip_ws();
      skip_comment();
      if (*p && *p != '\n') return error("expected end of line after value of " + key);
    }
  }
};

// ---------------------------------------------------------------------------
// schema helpers (serde semantics of ThetaJobToml: integers are accepted for
// f64 fields, floats are not accepted for integer fields)
// ---------------------------------------------------------------------------
struct Schema {
  std::string err;
  bool ok() const { return err.empty(); }
  void bad(const std::string& key, const char* want) {
    if (err.empty()) err = "invalid type for " + key + ": expected " + want;
  }
  bool get_f64(const Value& v, const std::string& k, double& out) {
    if (v.kind == Value::FLT) { out = v.f; return true; }
    if (v.kind == Value::INT) { out = (double)v.i; return true; }
    bad(k, "a number");
    return false;
  }
  bool get_usize(const Value& v, const std::string& k, int64_t& out) {
    if (v.kind == Value::INT && v.i >= 0) { out = v.i; return true; }
    bad(k, "a nonnegative integer");
    return false;
  }
  bool get_u64(const Value& v, const std::string& k, uint64_t& out) {
    if (v.kind == Value::INT && v.i >= 0) { out = (uint64_t)v.i; return true; }
    bad(k, "a nonnegative integer");
    return false;
  }
  bool get_vec_f64(const Value& v, const std::string& k, std::vector<double>& out) {
    if (v.kind != Value::ARR) { bad(k, "an array of numbers"); return false; }
    out.clear();
    for (const auto& e : v.a) {
      double d;
      if (!get_f64(e, k, d)) return false;
      out.push_back(d);
    }
    return true;
  }
  bool get_vec_usize(const Value& v, const std::string& k, std::vector<int64_t>& out) {
    if (v.kind != Value::ARR) { bad(k, "an array of nonnegative integers"); return false; }
    out.clear();
    for (const auto& e : v.a) {
      int64_t d;
      if (!get_usize(e, k, d)) return false;
      out.push_back(d);
    }
    return true;
  }
};

// Looks up `name` or any of its aliases; two spellings at once are an error
// like serde's "duplicate field".
const Value* lookup(const Table& t, std::initializer_list<const char*> names, Schema& sc) {
  const Value* found = nullptr;
  for (const char* n : names) {
    auto it = t.find(n);
    if (it == t.end()) continue;
    if (found && sc.ok()) sc.err = std::string("duplicate field `") + *names.begin() + "`";
    found = &it->second;
  }
  return found;
}

// parse_duration (src/model.rs:1033-1060), with Rust's ParseIntError texts
bool parse_u64_rust(const std::string& s, uint64_t& out, std::string& why) {
  if (s.empty()) { why = "cannot parse integer from empty string"; return false; }
  size_t i = (s[0] == '+') ? 1 : 0;
  if (i == s.size()) { why = "invalid digit found in string"; return false; }
  uint64_t v = 0;
  for (; i < s.size(); ++i) {
    if (s[i] < '0' || s[i] > '9') { why = "invalid digit found in string"; return false; }
    const uint64_t d = (uint64_t)(s[i] - '0');
    if (v > (UINT64_MAX - d) / 10) { why = "number too large to fit in target type"; return false; }
    v = v * 10 + d;
  }
  out = v;
  return true;
}

bool parse_duration(const std::string& value, int64_t& seconds, std::string& err) {
  std::string date_part = "0", time_part = value;
  const size_t dash = value.find('-');
  if (dash != std::string::npos) {
    date_part = value.substr(0, dash);
    time_part = value.substr(dash + 1);
  }
  uint64_t days;
  std::string why;
  if (!parse_u64_rust(date_part, days, why)) {
    err = "failed to parse duration days in " + value + ": " + why;
    return false;
  }
  std::vector<std::string> parts;
  size_t start = 0;
  for (;;) {
    const size_t c = time_part.find(':', start);
    parts.push_back(time_part.substr(start, c == std::string::npos ? std::string::npos : c - start));
    if (c == std::string::npos) break;
    start = c + 1;
  }
  if (parts.empty() || parts.size() > 3) {
    err = value + " does not match [[HH:]MM:]SS";
    return false;
  }
  std::vector<uint64_t> nums;
  for (const auto& part : parts) {
    uint64_t n;
    if (!parse_u64_rust(part, n, why)) {
      err = "failed to parse duration component in " + value + ": " + why;
      return false;
    }
    nums.push_back(n);
  }
  uint64_t h = 0, m = 0, s = 0;
  if (nums.size() == 1) s = nums[0];
  else if (nums.size() == 2) { m = nums[0]; s = nums[1]; }
  else { h = nums[0]; m = nums[1]; s = nums[2]; }
  seconds = (int64_t)(days * 86400 + h * 3600 + m * 60 + s);
  return true;
}

// Rust's f64::round: half away from zero; `as u64` saturates at 0
uint64_t round_u64(double x) {
  const double r = std::round(x);
  if (!(r > 0.0)) return 0;
  if (r >= 18446744073709551615.0) return UINT64_MAX;
  return (uint64_t)r;
}

}  // namespace

struct TaskRecord {
  slsf_task t;
  std::vector<double> j_xy, j_z;
};

struct slsf_job {
  std::string name;
  int64_t run_time_s = 12 * 3600, checkpoint_time_s = 30 * 60;
  std::vector<TaskRecord> tasks;
};

namespace {

// ThetaJobConfig with the reference defaults (src/model.rs:518-546)
struct JobConfig {
  std::vector<int64_t> l;
  bool has_lx = false, has_ly = false, has_lz = false;
  std::vector<int64_t> l_x, l_y, l_z;
  std::vector<double> temperatures{2.8, 2.9, 3.0, 3.1, 3.2, 3.3};
  std::vector<double> delta_j_z{0.0}, delta_j_xy{0.0};
  int64_t samples = 16;
  uint64_t base_seed = 20260414;
  double j_xy = 1.0, j_z_mean = 0.1;
  int64_t sweeps = 20000, thermalization = 5000, binsize = 50;
  double proposal_width = 3.141592653589793;
  int64_t wolff_steps = 1;
  bool has_rmax = false, has_rmax_xy = false, has_rmax_z = false;
  std::vector<int64_t> rmax, rmax_xy, rmax_z;
  int64_t correlation_interval = 1;
  int64_t run_time = 12 * 3600, checkpoint_time = 30 * 60;
  std::string job_name = "xy_theta_rust";
  // extension: uniform per-layer in-plane disorder
  int disorder_kind = 0;
  double j_xy_min = 0.0, j_xy_max = 0.0;
  JobConfig() {
    for (int64_t v = 4; v <= 20; ++v) l.push_back(v);
  }
};

int two_point(std::vector<double>& out, int64_t n, double mean, double delta, SplitMix64& rng,
              const char* name) {  // src/initialization.rs:5-31
  if (delta < 0.0) return fail(SLSF_E_INVALID, std::string("\xce\xb4") + name + " must be nonnegative");
  if (mean - delta < 0.0)
    return fail(SLSF_E_INVALID, std::string("two-point layer disorder requires ") + name +
                                    "_mean - \xce\xb4" + name + " >= 0");
  out.assign((size_t)n, mean);
  if (delta != 0.0)
    for (int64_t i = 0; i < n; ++i) out[i] = rng.next_bool() ? mean + delta : mean - delta;
  return SLSF_OK;
}

int make_tasks(const JobConfig& c, std::vector<TaskRecord>& tasks) {  // src/model.rs:697-810
  struct Spec { int64_t lx, ly, lz, l; };
  std::vector<Spec> specs;
  if (!c.has_lx && !c.has_ly && !c.has_lz) {
    for (int64_t v : c.l) specs.push_back({v, v, v, v});
  } else {
    const auto& xs = c.has_lx ? c.l_x : c.l;
    const auto& ys = c.has_ly ? c.l_y : c.l;
    const auto& zs = c.has_lz ? c.l_z : c.l;
    for (int64_t x : xs)
      for (int64_t y : ys)
        for (int64_t z : zs) specs.push_back({x, y, z, z});
  }
  if (specs.empty()) return fail(SLSF_E_INVALID, "at least one lattice size must be configured");
  const struct { const char* name; bool has; const std::vector<int64_t>* v; } counts[3] = {
      {"corr_rmax", c.has_rmax, &c.rmax},
      {"corr_rmax_xy", c.has_rmax_xy, &c.rmax_xy},
      {"corr_rmax_z", c.has_rmax_z, &c.rmax_z}};
  for (const auto& k : counts)
    if (k.has && k.v->size() != specs.size())
      return fail(SLSF_E_INVALID, std::string(k.name) +
                                      " must contain exactly one value per lattice size: expected " +
                                      std::to_string(specs.size()) + ", got " +
                                      std::to_string(k.v->size()));
  if (c.temperatures.empty()) return fail(SLSF_E_INVALID, "at least one temperature must be configured");
  if (c.delta_j_xy.empty() || c.delta_j_z.empty())
    return fail(SLSF_E_INVALID, "delta_j_xy and delta_j_z must not be empty");
  if (c.samples == 0) return fail(SLSF_E_INVALID, "samples must be positive");
  if (c.binsize == 0) return fail(SLSF_E_INVALID, "binsize must be positive");
  if (c.correlation_interval == 0) return fail(SLSF_E_INVALID, "correlation_interval must be positive");
  if (c.sweeps < c.binsize) return fail(SLSF_E_INVALID, "sweeps must be at least binsize");
  for (size_t li = 0; li < specs.size(); ++li) {
    const Spec& sp = specs[li];
    for (double djxy : c.delta_j_xy)
      for (double djz : c.delta_j_z)
        for (double temp : c.temperatures)
          for (int64_t sample = 1; sample <= c.samples; ++sample) {
            const uint64_t disorder_seed = c.base_seed + (uint64_t)sample - 1;
            const uint64_t seed = c.base_seed + 100000ull * (uint64_t)sp.lz +
                                  1000ull * (uint64_t)sample + round_u64(100.0 * temp) +
                                  round_u64(1000.0 * djxy) * 10000ull + round_u64(1000.0 * djz);
            SplitMix64 rng(disorder_seed);
            TaskRecord rec;
            if (c.disorder_kind == 1) {
              if (!(c.j_xy_max >= c.j_xy_min) || c.j_xy_min < 0.0)
                return fail(SLSF_E_INVALID, "uniform layer disorder requires 0 <= j_xy_min <= j_xy_max");
              rec.j_xy.resize((size_t)sp.lz);
              for (int64_t z = 0; z < sp.lz; ++z)
                rec.j_xy[z] = c.j_xy_min + (c.j_xy_max - c.j_xy_min) * rng.next_f64();
            } else if (int rc = two_point(rec.j_xy, sp.lz, c.j_xy, djxy, rng, "J_xy")) {
              return rc;
            }
            if (int rc = two_point(rec.j_z, sp.lz, c.j_z_mean, djz, rng, "J_z")) return rc;
            slsf_task& t = rec.t;
            std::memset(&t, 0, sizeof t);
            snprintf(t.name, sizeof t.name, "L%lldx%lldx%lld_T%.6f_dJxy%.6f_dJz%.6f_sample%lld",
                     (long long)sp.lx, (long long)sp.ly, (long long)sp.lz, temp, djxy, djz,
                     (long long)sample);
            t.l = (int32_t)sp.l; t.l_x = (int32_t)sp.lx; t.l_y = (int32_t)sp.ly; t.l_z = (int32_t)sp.lz;
            t.temperature = temp;
            t.j_xy = c.j_xy;
            t.delta_j_xy = djxy;
            t.j_z_mean = c.j_z_mean;
            t.delta_j_z = djz;
            t.disorder_seed = disorder_seed;
            t.seed = seed;
            t.sample = (int32_t)sample;
            t.sweeps = c.sweeps;
            t.thermalization = c.thermalization;
            t.binsize = (int32_t)c.binsize;
            t.proposal_width = c.proposal_width;
            t.wolff_steps = (int32_t)c.wolff_steps;
            t.correlation_rmax = (int32_t)(c.has_rmax ? c.rmax[li] : sp.lz / 2);
            t.correlation_rmax_xy = (int32_t)(c.has_rmax_xy ? c.rmax_xy[li] : std::min(sp.lx, sp.ly) / 2);
            t.correlation_rmax_z = (int32_t)(c.has_rmax_z ? c.rmax_z[li] : sp.lz / 2);
            t.correlation_interval = (int32_t)c.correlation_interval;
            t.disorder_kind = c.disorder_kind;
            tasks.push_back(std::move(rec));
          }
  }
  return SLSF_OK;
}

int config_from_doc(const std::map<std::string, Table>& doc, JobConfig& c) {
  Schema sc;
  const Table& root = doc.at("");
  if (auto* v = lookup(root, {"name"}, sc)) {
    if (v->kind != Value::STR) sc.bad("name", "a string");
    else c.job_name = v->s;
  }
  std::string derr;
  if (auto* v = lookup(root, {"run_time"}, sc)) {
    if (v->kind != Value::STR) sc.bad("run_time", "a string");
    else if (!parse_duration(v->s, c.run_time, derr))
      return fail(SLSF_E_INVALID, "invalid run_time: " + derr);
  }
  if (auto* v = lookup(root, {"checkpoint_time"}, sc)) {
    if (v->kind != Value::STR) sc.bad("checkpoint_time", "a string");
    else if (!parse_duration(v->s, c.checkpoint_time, derr))
      return fail(SLSF_E_INVALID, "invalid checkpoint_time: " + derr);
  }
  auto mit = doc.find("model");
  if (mit != doc.end()) {
    const Table& m = mit->second;
    if (auto* v = lookup(m, {"l", "L"}, sc)) sc.get_vec_usize(*v, "l", c.l);
    if (auto* v = lookup(m, {"l_x"}, sc)) c.has_lx = sc.get_vec_usize(*v, "l_x", c.l_x);
    if (auto* v = lookup(m, {"l_y"}, sc)) c.has_ly = sc.get_vec_usize(*v, "l_y", c.l_y);
    if (auto* v = lookup(m, {"l_z"}, sc)) c.has_lz = sc.get_vec_usize(*v, "l_z", c.l_z);
    // temperatures (alias T) take precedence over the separate `t` field
    const Value* tv = lookup(m, {"temperatures", "T"}, sc);
    if (!tv) tv = lookup(m, {"t"}, sc);
    if (tv) sc.get_vec_f64(*tv, "temperatures", c.temperatures);
    const Value* dz = lookup(m, {"delta_j_z"}, sc);
    if (!dz) dz = lookup(m, {"djz"}, sc);
    if (dz) sc.get_vec_f64(*dz, "delta_j_z", c.delta_j_z);
    const Value* dxy = lookup(m, {"delta_j_xy"}, sc);
    if (!dxy) dxy = lookup(m, {"djxy"}, sc);
    if (dxy) sc.get_vec_f64(*dxy, "delta_j_xy", c.delta_j_xy);
    if (auto* v = lookup(m, {"samples"}, sc)) sc.get_usize(*v, "samples", c.samples);
    if (auto* v = lookup(m, {"base_seed"}, sc)) sc.get_u64(*v, "base_seed", c.base_seed);
    if (auto* v = lookup(m, {"j_xy"}, sc)) sc.get_f64(*v, "j_xy", c.j_xy);
    const Value* jz = lookup(m, {"j_z_mean"}, sc);
    if (!jz) jz = lookup(m, {"j_z"}, sc);
    if (jz) sc.get_f64(*jz, "j_z_mean", c.j_z_mean);
    if (auto* v = lookup(m, {"disorder"}, sc)) {
      if (v->kind != Value::STR) sc.bad("disorder", "a string");
      else if (v->s == "two_point") c.disorder_kind = 0;
      else if (v->s == "uniform") c.disorder_kind = 1;
      else return fail(SLSF_E_INVALID, "disorder must be \"two_point\" or \"uniform\"");
    }
    if (auto* v = lookup(m, {"j_xy_min"}, sc)) sc.get_f64(*v, "j_xy_min", c.j_xy_min);
    if (auto* v = lookup(m, {"j_xy_max"}, sc)) sc.get_f64(*v, "j_xy_max", c.j_xy_max);
  }
  auto rit = doc.find("run");
  if (rit != doc.end()) {
    const Table& r = rit->second;
    if (auto* v = lookup(r, {"sweeps"}, sc)) sc.get_usize(*v, "sweeps", c.sweeps);
    if (auto* v = lookup(r, {"thermalization"}, sc)) sc.get_usize(*v, "thermalization", c.thermalization);
    if (auto* v = lookup(r, {"binsize"}, sc)) sc.get_usize(*v, "binsize", c.binsize);
    if (auto* v = lookup(r, {"proposal_width"}, sc)) sc.get_f64(*v, "proposal_width", c.proposal_width);
    if (auto* v = lookup(r, {"wolff_steps"}, sc)) sc.get_usize(*v, "wolff_steps", c.wolff_steps);
  }
  auto meit = doc.find("measure");
  if (meit != doc.end()) {  // src/model.rs:621-632
    const Table& me = meit->second;
    if (auto* v = lookup(me, {"corr_rmax"}, sc)) c.has_rmax = sc.get_vec_usize(*v, "corr_rmax", c.rmax);
    if (auto* v = lookup(me, {"corr_rmax_xy"}, sc)) c.has_rmax_xy = sc.get_vec_usize(*v, "corr_rmax_xy", c.rmax_xy);
    else if (c.has_rmax) { c.has_rmax_xy = true; c.rmax_xy = c.rmax; }
    if (auto* v = lookup(me, {"corr_rmax_z"}, sc)) c.has_rmax_z = sc.get_vec_usize(*v, "corr_rmax_z", c.rmax_z);
    else if (c.has_rmax) { c.has_rmax_z = true; c.rmax_z = c.rmax; }
    const Value* ci = lookup(me, {"correlation_interval"}, sc);
    if (!ci) ci = lookup(me, {"corr_interval"}, sc);
    if (ci) sc.get_usize(*ci, "correlation_interval", c.correlation_interval);
  }
  if (!sc.ok()) return fail(SLSF_E_INVALID, sc.err);
  return SLSF_OK;
}

}  // namespace

extern "C" {

int slsf_parse_duration(const char* text, int64_t* seconds) {
  clear_error();
  if (!text || !seconds) return fail(SLSF_E_INVALID, "null argument");
  std::string err;
  if (!parse_duration(text, *seconds, err)) return fail(SLSF_E_INVALID, err);
  return SLSF_OK;
}

int slsf_job_from_toml(const char* toml_text, slsf_job** out) {
  clear_error();
  if (!toml_text || !out) return fail(SLSF_E_INVALID, "null argument");
  *out = nullptr;
  Parser ps{toml_text};
  std::map<std::string, Table> doc;
  if (!ps.parse(doc)) return fail(SLSF_E_INVALID, ps.err);
  JobConfig c;
  if (int rc = config_from_doc(doc, c)) return rc;
  auto job = std::make_unique<slsf_job>();
  job->name = c.job_name;
  job->run_time_s = c.run_time;
  job->checkpoint_time_s = c.checkpoint_time;
  if (int rc = make_tasks(c, job->tasks)) return rc;
  *out = job.release();
  return SLSF_OK;
}

void slsf_job_destroy(slsf_job* job) { delete job; }
const char* slsf_job_name(const slsf_job* job) { return job ? job->name.c_str() : ""; }
int64_t slsf_job_num_tasks(const slsf_job* job) { return job ? (int64_t)job->tasks.size() : 0; }

int slsf_job_task(const slsf_job* job, int64_t index, slsf_task* out) {
  clear_error();
  if (!job || !out) return fail(SLSF_E_INVALID, "null argument");
  if (index < 0 || index >= (int64_t)job->tasks.size())
    return fail(SLSF_E_INVALID, "task index out of range");
  *out = job->tasks[index].t;
  return SLSF_OK;
}

int slsf_job_task_couplings(const slsf_job* job, int64_t index, double* j_xy, double* j_z) {
  clear_error();
  if (!job) return fail(SLSF_E_INVALID, "null argument");
  if (index < 0 || index >= (int64_t)job->tasks.size())
    return fail(SLSF_E_INVALID, "task index out of range");
  const auto& rec = job->tasks[index];
  if (j_xy) std::memcpy(j_xy, rec.j_xy.data(), sizeof(double) * rec.j_xy.size());
  if (j_z) std::memcpy(j_z, rec.j_z.data(), sizeof(double) * rec.j_z.size());
  return SLSF_OK;
}

int slsf_job_durations(const slsf_job* job, int64_t* run_time_s, int64_t* checkpoint_time_s) {
  clear_error();
  if (!job) return fail(SLSF_E_INVALID, "null argument");
  if (run_time_s) *run_time_s = job->run_time_s;
  if (checkpoint_time_s) *checkpoint_time_s = job->checkpoint_time_s;
  return SLSF_OK;
}

int64_t slsf_job_selected_tasks(const slsf_job* job, int32_t rank, int32_t world, int64_t* indices,
                                int64_t cap) {
  if (!job || world <= 0 || rank < 0 || rank >= world) return -1;
  int64_t n = 0;
  for (int64_t i = 0; i < (int64_t)job->tasks.size(); ++i)
    if (world == 1 || i % world == rank) {
      if (indices && n < cap) indices[n] = i;
      ++n;
    }
  return n;
}

}  // extern "C"
