// Calibration profiles: measured launch settings (the autotuner's output)
// applied when a context is created, and the build identity they are
// matched against.
//
// A profile is JSON written by dataframe_mojo_amd/autotune.py:
//   {"version": 2,
//    "identity": {"device": "...", "arch": "gfx950", "compute_units": 256,
//                 "build_id": "<this library's DFR_BUILD_ID>"},
//    "tuning": {"scan_block": 1024, ...},          (dfr_context_tune keys)
//    "transfer": {...}, "groupby": {...}, ...}     (placement costs, read by
//                                                   the host provider)
// It is found at $DATAFRAME_ROCM_CALIBRATION (a file; "off" disables it) or
// at calibration/<arch>-<CUs>cu.json beside libdfrocm.so. A profile is
// applied only when every identity field matches the device and this
// library's build: settings measured on other hardware or by other code are
// never used (as the reference's versioned, identity-matched profiles,
// dataframe_metal/calibration.mojo:1-12, whose unidentified or legacy
// profiles carry no measured cost).
#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>

#include "../../include/dfrocm.h"
#include "runtime.h"

#ifndef DFR_BUILD_ID
#define DFR_BUILD_ID "unidentified"
#endif

namespace {

constexpr int kCalibrationVersion = 2;

// A JSON value reduced to what a profile needs: objects (by key), strings
// and numbers; arrays, booleans and null are parsed and kept as their kind.
struct Json {
  enum Kind { NUL, NUM, STR, OBJ, OTHER } kind = NUL;
  double num = 0;
  std::string str;
  std::map<std::string, Json> obj;
  const Json *get(const char *key) const {
    if (kind != OBJ) return nullptr;
    auto it = obj.find(key);
    return it == obj.end() ? nullptr : &it->second;
  }
};

struct Parser {
  const char *p, *end;
  bool ok = true;
  void ws() {
    while (p < end && isspace((unsigned char)*p)) ++p;
  }
  bool lit(const char *s) {
    size_t n = strlen(s);
    if ((size_t)(end - p) < n || strncmp(p, s, n) != 0) return false;
    p += n;
    return true;
  }
  std::string string() {
    std::string out;
    ++p;  // opening quote
    while (p < end && *p != '"') {
      if (*p == '\\' && p + 1 < end) {
        ++p;
        switch (*p) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'u':
            // \uXXXX: kept as '?' (profiles hold ASCII)
            p += std::min<ptrdiff_t>(4, end - p - 1);
            out += '?';
            break;
          default: out += *p;
        }
        ++p;
        continue;
      }
      out += *p++;
    }
    if (p >= end) ok = false;
    ++p;  // closing quote
    return out;
  }
  Json value(int depth) {
    Json v;
    ws();
    if (p >= end || depth > 32) {
      ok = false;
      return v;
    }
    if (*p == '{') {
      v.kind = Json::OBJ;
      ++p;
      ws();
      if (p < end && *p == '}') {
        ++p;
        return v;
      }
      while (ok) {
        ws();
        if (p >= end || *p != '"') {
          ok = false;
          break;
        }
        std::string k = string();
        ws();
        if (p >= end || *p != ':') {
          ok = false;
          break;
        }
        ++p;
        v.obj[k] = value(depth + 1);
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        if (p < end && *p == '}') {
          ++p;
          break;
        }
        ok = false;
      }
      return v;
    }
    if (*p == '[') {
      v.kind = Json::OTHER;
      ++p;
      ws();
      if (p < end && *p == ']') {
        ++p;
        return v;
      }
      while (ok) {
        value(depth + 1);
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        if (p < end && *p == ']') {
          ++p;
          break;
        }
        ok = false;
      }
      return v;
    }
    if (*p == '"') {
      v.kind = Json::STR;
      v.str = string();
      return v;
    }
    if (lit("true") || lit("false")) {
      v.kind = Json::OTHER;
      return v;
    }
    if (lit("null")) return v;
    char *stop = nullptr;
    v.num = strtod(p, &stop);
    if (stop == p) {
      ok = false;
      return v;
    }
    v.kind = Json::NUM;
    p = stop;
    return v;
  }
};

bool read_all(const std::string &path, std::string &out) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) return false;
  char buf[65536];
  size_t n;
  out.clear();
  while ((n = fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, n);
  fclose(f);
  return true;
}

// The directory libdfrocm.so was loaded from.
std::string library_dir() {
  Dl_info info;
  if (!dladdr((const void *)&dfr_abi_version, &info) || !info.dli_fname) return ".";
  std::string p(info.dli_fname);
  size_t slash = p.rfind('/');
  return slash == std::string::npos ? std::string(".") : p.substr(0, slash);
}

std::string arch_of(const hipDeviceProp_t &prop) {
  std::string a(prop.gcnArchName);
  size_t colon = a.find(':');
  return colon == std::string::npos ? a : a.substr(0, colon);
}

}  // namespace

extern "C" const char *dfr_build_id(void) { return DFR_BUILD_ID; }

// Applies the device's calibration profile to a new context's tuning and
// records what happened in ctx->calibration (never fails the context: a
// missing or mismatched profile leaves the built-in settings).
void apply_calibration(dfr_context *ctx, const hipDeviceProp_t &prop, const char *device_name) {
  const std::string arch = arch_of(prop);
  std::string path;
  const char *env = getenv("DATAFRAME_ROCM_CALIBRATION");
  if (env && *env) {
    if (strcmp(env, "off") == 0) {
      ctx->calibration = "none: DATAFRAME_ROCM_CALIBRATION=off";
      return;
    }
    path = env;
  } else {
    path = library_dir() + "/calibration/" + arch + "-" + std::to_string(prop.multiProcessorCount) + "cu.json";
  }
  std::string text;
  if (!read_all(path, text)) {
    ctx->calibration = "none: no profile at " + path;
    return;
  }
  Parser ps{text.data(), text.data() + text.size()};
  Json root = ps.value(0);
  if (!ps.ok || root.kind != Json::OBJ) {
    ctx->calibration = "none: " + path + " is not a JSON object";
    return;
  }
  const Json *ver = root.get("version");
  if (!ver || ver->kind != Json::NUM || (int)ver->num != kCalibrationVersion) {
    ctx->calibration = "none: " + path + " has another profile version";
    return;
  }
  const Json *id = root.get("identity");
  auto id_str = [&](const char *k) -> std::string {
    const Json *v = id ? id->get(k) : nullptr;
    return v && v->kind == Json::STR ? v->str : std::string();
  };
  const Json *cu = id ? id->get("compute_units") : nullptr;
  std::string why;
  if (!id || id->kind != Json::OBJ)
    why = "no identity";
  else if (id_str("arch") != arch)
    why = "arch " + id_str("arch") + " != " + arch;
  else if (!cu || cu->kind != Json::NUM || (int)cu->num != prop.multiProcessorCount)
    why = "compute units differ";
  else if (id_str("device") != device_name)
    why = "device " + id_str("device") + " != " + device_name;
  else if (id_str("build_id") != DFR_BUILD_ID)
    why = "measured with build " + id_str("build_id") + ", this is " + DFR_BUILD_ID;
  if (!why.empty()) {
    ctx->calibration = "none: " + path + " does not match (" + why + ")";
    return;
  }
  int applied = 0, refused = 0;
  if (const Json *tune = root.get("tuning")) {
    for (auto &kv : tune->obj) {
      if (kv.second.kind != Json::NUM) {
        ++refused;
        continue;
      }
      char *err = nullptr;
      if (dfr_context_tune(ctx, kv.first.c_str(), (int64_t)kv.second.num, &err) == DFR_OK) {
        ++applied;
      } else {
        ++refused;
        free(err);
      }
    }
  }
  ctx->calibration = "applied: " + path + " (" + std::to_string(applied) + " settings" +
                     (refused ? ", " + std::to_string(refused) + " refused" : std::string()) + ")";
}

extern "C" int dfr_context_calibration(const dfr_context *ctx, char **text, char **error) {
  if (!ctx || !text) return dfr_fail(error, DFR_INVALID, "null argument");
  *text = strdup(ctx->calibration.c_str());
  return *text ? DFR_OK : dfr_fail(error, DFR_DEVICE, "out of host memory");
}
