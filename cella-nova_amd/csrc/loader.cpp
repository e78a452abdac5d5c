// Creature loader: the C ABI counterpart of ml-rs/config.rs (SimConfig, load_config,
// load_kernels).  No JSON library exists in this image, so a small recursive-descent reader
// is used; it mirrors serde's behaviour for the SimConfig fields (unknown keys ignored,
// missing keys rejected, unsigned fields must be non-negative integers).
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "common.h"

namespace cn {
namespace {

struct Reader {
  const char* p;
  const char* end;
  std::string err;

  void ws() {
    while (p < end && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) ++p;
  }
  bool fail(const char* msg) {
    if (err.empty()) {
      char buf[256];
      snprintf(buf, sizeof buf, "%s at byte %ld", msg, (long)(p - start));
      err = buf;
    }
    return false;
  }
  const char* start;
  bool expect(char c) {
    ws();
    if (p >= end || *p != c) {
      char m[64];
      snprintf(m, sizeof m, "expected '%c'", c);
      return fail(m);
    }
    ++p;
    return true;
  }
  bool string(std::string& out) {
    ws();
    if (p >= end || *p != '"') return fail("expected string");
    ++p;
    out.clear();
    while (p < end && *p != '"') {
      if (*p == '\\') {
        ++p;
        if (p >= end) return fail("bad escape");
        if (*p == 'u') {
          if (end - p < 5) return fail("bad \\u escape");
          p += 4;  // keys of interest are ASCII; keep a placeholder
          out.push_back('?');
        } else {
          out.push_back(*p);
        }
        ++p;
      } else {
        out.push_back(*p++);
      }
    }
    if (p >= end) return fail("unterminated string");
    ++p;
    return true;
  }
  // Number token; is_int reports whether it had no fraction or exponent.
  bool number(double& v, bool& is_int) {
    ws();
    const char* s = p;
    if (p < end && (*p == '-' || *p == '+')) ++p;
    bool digits = false;
    is_int = true;
    while (p < end && ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E' || *p == '-' || *p == '+')) {
      if (*p >= '0' && *p <= '9') digits = true;
      if (*p == '.' || *p == 'e' || *p == 'E') is_int = false;
      ++p;
    }
    if (!digits) {
      p = s;
      return fail("expected number");
    }
    std::string tok(s, p);
    char* e = nullptr;
    errno = 0;
    v = std::strtod(tok.c_str(), &e);
    if (e != tok.c_str() + tok.size()) return fail("malformed number");
    if (!std::isfinite(v)) return fail("number out of range");  // as serde_json: no f64 overflow
    return true;
  }
  bool skip_value(int depth = 0) {
    if (depth > 64) return fail("nesting too deep");
    ws();
    if (p >= end) return fail("unexpected end");
    if (*p == '{') {
      ++p;
      ws();
      if (p < end && *p == '}') {
        ++p;
        return true;
      }
      for (;;) {
        std::string k;
        if (!string(k) || !expect(':') || !skip_value(depth + 1)) return false;
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        return expect('}');
      }
    }
    if (*p == '[') {
      ++p;
      ws();
      if (p < end && *p == ']') {
        ++p;
        return true;
      }
      for (;;) {
        if (!skip_value(depth + 1)) return false;
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        return expect(']');
      }
    }
    if (*p == '"') {
      std::string s;
      return string(s);
    }
    if (end - p >= 4 && (!strncmp(p, "true", 4) || !strncmp(p, "null", 4))) {
      p += 4;
      return true;
    }
    if (end - p >= 5 && !strncmp(p, "false", 5)) {
      p += 5;
      return true;
    }
    double v;
    bool ii;
    return number(v, ii);
  }
  template <class F>
  bool array(F&& elem) {
    if (!expect('[')) return false;
    ws();
    if (p < end && *p == ']') {
      ++p;
      return true;
    }
    for (;;) {
      if (!elem()) return false;
      ws();
      if (p < end && *p == ',') {
        ++p;
        continue;
      }
      return expect(']');
    }
  }
  bool uint_value(uint64_t& out) {
    double v;
    bool ii;
    if (!number(v, ii)) return false;
    if (!ii || v < 0 || v > 4294967295.0) return fail("expected a non-negative integer");
    out = (uint64_t)v;
    return true;
  }
  bool f64_array(std::vector<double>& out) {
    out.clear();
    return array([&] {
      double v;
      bool ii;
      if (!number(v, ii)) return false;
      out.push_back(v);
      return true;
    });
  }
  bool u32_array(std::vector<uint32_t>& out) {
    out.clear();
    return array([&] {
      uint64_t v;
      if (!uint_value(v)) return false;
      out.push_back((uint32_t)v);
      return true;
    });
  }
};

bool read_file(const char* path, std::vector<char>& buf) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  const long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (n < 0) {
    std::fclose(f);
    return false;
  }
  buf.resize((size_t)n);
  const size_t got = n ? std::fread(buf.data(), 1, (size_t)n, f) : 0;
  std::fclose(f);
  return got == (size_t)n;
}

}  // namespace
}  // namespace cn

using namespace cn;

extern "C" {

void cn_free_config(cn_config* cfg) {
  if (!cfg) return;
  std::free(cfg->seed_channels);
  std::free(cfg->c0);
  std::free(cfg->c1);
  std::free(cfg->growth_mu);
  std::free(cfg->growth_sigma);
  std::free(cfg->growth_weights);
  std::free(cfg);
}

cn_status cn_load_config(const char* path, cn_config** out) {
  CN_TRY
  clear_error();
  if (!path || !out) return set_error(CN_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  std::vector<char> buf;
  if (!read_file(path, buf)) return set_error(CN_ERR_IO, "cannot read config file '%s'", path);
  Reader r{buf.data(), buf.data() + buf.size(), {}, buf.data()};
  uint64_t seed_size = 0, nc = 0, nk = 0;
  bool has[9] = {false};
  std::vector<std::vector<double>> seeds;
  std::vector<uint32_t> c0, c1;
  std::vector<double> mu, sigma, w;
  bool ok = r.expect('{');
  if (ok) {
    r.ws();
    if (r.p < r.end && *r.p == '}') {
      ++r.p;
    } else {
      for (;;) {
        std::string key;
        if (!(ok = r.string(key) && r.expect(':'))) break;
        if (key == "seed_size") ok = r.uint_value(seed_size), has[0] = true;
        else if (key == "num_channels") ok = r.uint_value(nc), has[1] = true;
        else if (key == "num_kernels") ok = r.uint_value(nk), has[2] = true;
        else if (key == "seed_channels") {
          has[3] = true;
          seeds.clear();
          ok = r.array([&] {
            seeds.emplace_back();
            return r.f64_array(seeds.back());
          });
        } else if (key == "c0") ok = r.u32_array(c0), has[4] = true;
        else if (key == "c1") ok = r.u32_array(c1), has[5] = true;
        else if (key == "growth_mu") ok = r.f64_array(mu), has[6] = true;
        else if (key == "growth_sigma") ok = r.f64_array(sigma), has[7] = true;
        else if (key == "growth_weights") ok = r.f64_array(w), has[8] = true;
        else ok = r.skip_value();
        if (!ok) break;
        r.ws();
        if (r.p < r.end && *r.p == ',') {
          ++r.p;
          continue;
        }
        ok = r.expect('}');
        break;
      }
    }
  }
  if (!ok) return set_error(CN_ERR_PARSE, "config '%s': %s", path, r.err.c_str());
  r.ws();
  if (r.p != r.end) return set_error(CN_ERR_PARSE, "config '%s': trailing characters", path);
  static const char* names[9] = {"seed_size", "num_channels", "num_kernels", "seed_channels", "c0",
                                 "c1", "growth_mu", "growth_sigma", "growth_weights"};
  for (int i = 0; i < 9; ++i)
    if (!has[i]) return set_error(CN_ERR_PARSE, "config '%s': missing field `%s`", path, names[i]);
  if (c0.size() != nk || c1.size() != nk || mu.size() != nk || sigma.size() != nk || w.size() != nk)
    return set_error(CN_ERR_PARSE, "config '%s': c0/c1/growth arrays must have num_kernels = %llu entries", path,
                     (unsigned long long)nk);
  if (seeds.size() != nc)
    return set_error(CN_ERR_PARSE, "config '%s': %zu seed channels, num_channels = %llu", path, seeds.size(),
                     (unsigned long long)nc);
  const size_t P = (size_t)seed_size * seed_size;
  for (size_t c = 0; c < seeds.size(); ++c)
    if (seeds[c].size() != P)
      return set_error(CN_ERR_PARSE, "config '%s': seed channel %zu has %zu values, expected seed_size^2 = %zu",
                       path, c, seeds[c].size(), P);

  cn_config* cfg = (cn_config*)std::calloc(1, sizeof(cn_config));
  if (!cfg) return set_error(CN_ERR_OUT_OF_MEMORY, "host allocation failed");
  cfg->seed_size = (uint32_t)seed_size;
  cfg->num_channels = (uint32_t)nc;
  cfg->num_kernels = (uint32_t)nk;
  cfg->seed_len = P;
  cfg->seed_channels = (double*)std::malloc(std::max<size_t>(1, nc * P) * sizeof(double));
  cfg->c0 = (uint32_t*)std::malloc(std::max<size_t>(1, nk) * 4);
  cfg->c1 = (uint32_t*)std::malloc(std::max<size_t>(1, nk) * 4);
  cfg->growth_mu = (float*)std::malloc(std::max<size_t>(1, nk) * 4);
  cfg->growth_sigma = (float*)std::malloc(std::max<size_t>(1, nk) * 4);
  cfg->growth_weights = (float*)std::malloc(std::max<size_t>(1, nk) * 4);
  if (!cfg->seed_channels || !cfg->c0 || !cfg->c1 || !cfg->growth_mu || !cfg->growth_sigma || !cfg->growth_weights) {
    cn_free_config(cfg);
    return set_error(CN_ERR_OUT_OF_MEMORY, "host allocation failed");
  }
  for (size_t c = 0; c < nc; ++c) std::memcpy(cfg->seed_channels + c * P, seeds[c].data(), P * sizeof(double));
  for (size_t k = 0; k < nk; ++k) {
    // config.rs:40 swaps c0 and c1 so that c0 is the input channel in Python's convention.
    cfg->c0[k] = c1[k];
    cfg->c1[k] = c0[k];
    cfg->growth_mu[k] = (float)mu[k];
    cfg->growth_sigma[k] = (float)sigma[k];
    cfg->growth_weights[k] = (float)w[k];
  }
  *out = cfg;
  return CN_OK;
  CN_CATCH
}

cn_status cn_load_kernels(const char* path, uint32_t num_kernels, uint32_t target, float* out, size_t n_floats) {
  CN_TRY
  clear_error();
  if (!path || !out) return set_error(CN_ERR_INVALID_ARG, "null argument");
  if (num_kernels == 0 || target == 0) return set_error(CN_ERR_INVALID_ARG, "num_kernels and target must be > 0");
  const size_t T = target, TT = T * T;
  if (n_floats != 2 * (size_t)num_kernels * TT)
    return set_error(CN_ERR_INVALID_ARG, "output has %zu floats, expected %zu", n_floats, 2 * (size_t)num_kernels * TT);
  std::vector<char> buf;
  if (!read_file(path, buf)) return set_error(CN_ERR_IO, "cannot read kernel file '%s'", path);
  const size_t fs = buf.size();
  if (fs == 0 || fs % ((size_t)num_kernels * 8) != 0)
    return set_error(CN_ERR_IO, "kernel file '%s' has %zu bytes, not a multiple of K*8 = %zu", path, fs,
                     (size_t)num_kernels * 8);
  const size_t elems = fs / num_kernels / 8;
  size_t S = (size_t)std::llround(std::sqrt((double)elems));
  if (S * S != elems)
    return set_error(CN_ERR_IO, "kernel file '%s': %zu complex values per kernel is not a square", path, elems);
  if (S > T) return set_error(CN_ERR_INVALID_ARG, "stored kernel size %zu exceeds target %zu", S, T);
  if ((T - S) % 2 != 0) return set_error(CN_ERR_UNSUPPORTED, "target %zu - stored %zu must be even", T, S);
  auto rd = [&](size_t k, size_t i, float* re, float* im) {
    const char* b = buf.data() + (k * S * S + i) * 8;
    std::memcpy(re, b, 4);  // little-endian f32 on every supported host
    std::memcpy(im, b + 4, 4);
  };
  for (size_t k = 0; k < num_kernels; ++k) {
    float* o = out + k * TT * 2;
    if (S == T) {
      for (size_t i = 0; i < TT; ++i) rd(k, i, &o[2 * i], &o[2 * i + 1]);
      continue;
    }
    // Spectral zero-padding (config.rs:81-117): fftshift the S*S spectrum, place it at the
    // centre of a T*T zero grid, then shift back by T/2.
    std::vector<float> padded(TT * 2, 0.f);
    const size_t hs = S / 2, pad = (T - S) / 2;
    for (size_t i = 0; i < S; ++i)
      for (size_t j = 0; j < S; ++j) {
        const size_t ni = (i + hs) % S, nj = (j + hs) % S;
        float re, im;
        rd(k, i * S + j, &re, &im);
        const size_t dst = ((pad + ni) * T + (pad + nj)) * 2;
        padded[dst] = re;
        padded[dst + 1] = im;
      }
    const size_t hd = T / 2;
    for (size_t i = 0; i < T; ++i)
      for (size_t j = 0; j < T; ++j) {
        const size_t ni = (i + hd) % T, nj = (j + hd) % T;
        o[(ni * T + nj) * 2] = padded[(i * T + j) * 2];
        o[(ni * T + nj) * 2 + 1] = padded[(i * T + j) * 2 + 1];
      }
  }
  return CN_OK;
  CN_CATCH
}

cn_status cn_write_kernels(const char* path, const float* kernels_fft, uint32_t num_kernels, uint32_t n) {
  CN_TRY
  clear_error();
  if (!path || !kernels_fft) return set_error(CN_ERR_INVALID_ARG, "null argument");
  if (num_kernels == 0 || n == 0) return set_error(CN_ERR_INVALID_ARG, "num_kernels and n must be > 0");
  FILE* f = std::fopen(path, "wb");
  if (!f) return set_error(CN_ERR_IO, "cannot open '%s' for writing: %s", path, std::strerror(errno));
  const size_t count = 2 * (size_t)num_kernels * n * n;  // little-endian f32 on every supported host
  const size_t wrote = std::fwrite(kernels_fft, sizeof(float), count, f);
  const bool closed = std::fclose(f) == 0;
  if (wrote != count || !closed) return set_error(CN_ERR_IO, "write to '%s' failed", path);
  return CN_OK;
  CN_CATCH
}

cn_status cn_write_config(const char* path, const cn_config* cfg) {
  CN_TRY
  clear_error();
  if (!path || !cfg) return set_error(CN_ERR_INVALID_ARG, "null argument");
  const size_t C = cfg->num_channels, K = cfg->num_kernels, P = (size_t)cfg->seed_size * cfg->seed_size;
  if (cfg->seed_len != P) return set_error(CN_ERR_INVALID_ARG, "seed_len %zu != seed_size^2 = %zu", cfg->seed_len, P);
  if ((C && P && !cfg->seed_channels) || (K && (!cfg->c0 || !cfg->c1 || !cfg->growth_mu || !cfg->growth_sigma ||
                                               !cfg->growth_weights)))
    return set_error(CN_ERR_INVALID_ARG, "null array in config");
  for (size_t j = 0; j < C * P; ++j)
    if (!std::isfinite(cfg->seed_channels[j])) return set_error(CN_ERR_INVALID_ARG, "seed value %zu not finite", j);
  std::string o;
  o.reserve(C * P * 24 + K * 64 + 256);
  char num[40];
  auto put_u = [&](const uint32_t* a) {
    o += '[';
    for (size_t k = 0; k < K; ++k) {
      std::snprintf(num, sizeof num, k ? ", %u" : "%u", a[k]);
      o += num;
    }
    o += ']';
  };
  auto put_f = [&](const float* a) {  // 9 significant digits: reads back to the same f32
    o += '[';
    for (size_t k = 0; k < K; ++k) {
      std::snprintf(num, sizeof num, k ? ", %.9g" : "%.9g", (double)a[k]);
      o += num;
    }
    o += ']';
  };
  std::snprintf(num, sizeof num, "%u", cfg->seed_size);
  o += "{\n  \"seed_size\": ";
  o += num;
  std::snprintf(num, sizeof num, "%u", cfg->num_channels);
  o += ",\n  \"num_channels\": ";
  o += num;
  std::snprintf(num, sizeof num, "%u", cfg->num_kernels);
  o += ",\n  \"num_kernels\": ";
  o += num;
  o += ",\n  \"seed_channels\": [";
  for (size_t c = 0; c < C; ++c) {
    o += c ? ",\n    [" : "\n    [";
    for (size_t j = 0; j < P; ++j) {
      std::snprintf(num, sizeof num, j ? ", %.17g" : "%.17g", cfg->seed_channels[c * P + j]);
      o += num;
    }
    o += ']';
  }
  o += C ? "\n  ]" : "]";
  // The loader swaps c0 and c1 (config.rs:40); write the JSON convention so it reads back cfg.
  o += ",\n  \"c0\": ";
  put_u(cfg->c1);
  o += ",\n  \"c1\": ";
  put_u(cfg->c0);
  o += ",\n  \"growth_mu\": ";
  put_f(cfg->growth_mu);
  o += ",\n  \"growth_sigma\": ";
  put_f(cfg->growth_sigma);
  o += ",\n  \"growth_weights\": ";
  put_f(cfg->growth_weights);
  o += "\n}\n";
  FILE* f = std::fopen(path, "wb");
  if (!f) return set_error(CN_ERR_IO, "cannot open '%s' for writing: %s", path, std::strerror(errno));
  const size_t wrote = std::fwrite(o.data(), 1, o.size(), f);
  const bool closed = std::fclose(f) == 0;
  if (wrote != o.size() || !closed) return set_error(CN_ERR_IO, "write to '%s' failed", path);
  return CN_OK;
  CN_CATCH
}

}  // extern "C"
