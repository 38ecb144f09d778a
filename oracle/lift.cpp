// Oracle (test infrastructure only): nd_lift restated from the reference.
//   kernels:   crates/ndic-lift/src/kernel.rs:66-335
//   chunk:     crates/ndic-lift/src/chunk.rs:30-302
//   config:    crates/ndic-lift/src/lib.rs:122-242
//   bytes:     crates/ndic-zarr/src/lift.rs:72-346
#include <algorithm>
#include <cstring>
#include <sstream>

#include "oracle.hpp"

namespace orc {

void fail(int kind, const std::string& message, size_t offset) { throw Error{kind, offset, message}; }

size_t dtype_size(int d) {
  switch (d) {
    case kU8: case kI8: return 1;
    case kU16: case kI16: return 2;
    case kU32: case kI32: case kF32: return 4;
    case kU64: case kI64: case kF64: return 8;
  }
  fail(kInvalidArgument, "unknown dtype code");
}
bool dtype_signed(int d) { return d == kI8 || d == kI16 || d == kI32 || d == kI64 || d == kF32 || d == kF64; }
const char* dtype_name(int d) {
  static const char* n[] = {"u8", "i8", "u16", "i16", "u32", "i32", "u64", "i64", "f32", "f64"};
  return (d >= 0 && d <= 9) ? n[d] : "?";
}

namespace {

typedef __int128 i128;

std::string i128_str(i128 v) {
  if (v == 0) return "0";
  bool neg = v < 0;
  unsigned __int128 u = neg ? (unsigned __int128)(-(v + 1)) + 1 : (unsigned __int128)v;
  std::string s;
  while (u) { s.push_back(char('0' + int(u % 10))); u /= 10; }
  if (neg) s.push_back('-');
  std::reverse(s.begin(), s.end());
  return s;
}

template <class T> struct Plane;
template <> struct Plane<int32_t> { typedef uint32_t U; static constexpr int BITS = 32; };
template <> struct Plane<int64_t> { typedef uint64_t U; static constexpr int BITS = 64; };

template <class T> inline T wadd(T a, T b) { typedef typename Plane<T>::U U; return (T)((U)a + (U)b); }
template <class T> inline T wsub(T a, T b) { typedef typename Plane<T>::U U; return (T)((U)a - (U)b); }

// kernel.rs:133-149
template <class T> void delta_forward(T* seg, size_t n, size_t inner) {
  for (size_t k = n; k-- > 1;)
    for (size_t j = 0; j < inner; ++j) seg[k * inner + j] -= seg[(k - 1) * inner + j];
}
template <class T> void delta_inverse(T* seg, size_t n, size_t inner) {
  for (size_t k = 1; k < n; ++k)
    for (size_t j = 0; j < inner; ++j) seg[k * inner + j] = wadd(seg[k * inner + j], seg[(k - 1) * inner + j]);
}

// kernel.rs:167-183
template <class T> void haar_forward_level(T* x, size_t n, size_t inner, T* tmp) {
  size_t ns = (n + 1) / 2;
  T* ts = tmp;
  T* td = tmp + ns * inner;
  for (size_t i = 0; i < n / 2; ++i)
    for (size_t j = 0; j < inner; ++j) {
      T x0 = x[2 * i * inner + j], x1 = x[(2 * i + 1) * inner + j];
      T d = x1 - x0;
      td[i * inner + j] = d;
      ts[i * inner + j] = x0 + (d >> 1);
    }
  if (n % 2 == 1)
    for (size_t j = 0; j < inner; ++j) ts[(ns - 1) * inner + j] = x[(n - 1) * inner + j];
  std::memcpy(x, tmp, n * inner * sizeof(T));
}
// kernel.rs:186-200
template <class T> void haar_inverse_level(T* x, size_t n, size_t inner, T* tmp) {
  size_t ns = (n + 1) / 2;
  const T* s = x;
  const T* d = x + ns * inner;
  for (size_t i = 0; i < n / 2; ++i)
    for (size_t j = 0; j < inner; ++j) {
      T x0 = wsub(s[i * inner + j], (T)(d[i * inner + j] >> 1));
      tmp[2 * i * inner + j] = x0;
      tmp[(2 * i + 1) * inner + j] = wadd(x0, d[i * inner + j]);
    }
  if (n % 2 == 1)
    for (size_t j = 0; j < inner; ++j) tmp[(n - 1) * inner + j] = s[(ns - 1) * inner + j];
  std::memcpy(x, tmp, n * inner * sizeof(T));
}

// kernel.rs:203-205
inline size_t right_even(size_t k, size_t n) { return k < n ? k : 2 * n - 2 - k; }

// kernel.rs:214-240
template <class T> void lift53_forward_level(T* x, size_t n, size_t inner, T* tmp) {
  size_t ns = (n + 1) / 2, nd = n / 2;
  T* ts = tmp;
  T* td = tmp + ns * inner;
  for (size_t i = 0; i < nd; ++i) {
    size_t r = right_even(2 * i + 2, n);
    for (size_t j = 0; j < inner; ++j)
      td[i * inner + j] = x[(2 * i + 1) * inner + j] - ((x[2 * i * inner + j] + x[r * inner + j] + 1) >> 1);
  }
  for (size_t i = 0; i < ns; ++i) {
    size_t il = i == 0 ? 0 : i - 1, ir = std::min(i, nd - 1);
    for (size_t j = 0; j < inner; ++j)
      ts[i * inner + j] = x[2 * i * inner + j] + ((td[il * inner + j] + td[ir * inner + j] + 2) >> 2);
  }
  std::memcpy(x, tmp, n * inner * sizeof(T));
}
// kernel.rs:243-272
template <class T> void lift53_inverse_level(T* x, size_t n, size_t inner, T* tmp) {
  size_t ns = (n + 1) / 2, nd = n / 2;
  const T* s = x;
  const T* d = x + ns * inner;
  for (size_t i = 0; i < ns; ++i) {
    size_t il = i == 0 ? 0 : i - 1, ir = std::min(i, nd - 1);
    for (size_t j = 0; j < inner; ++j)
      tmp[2 * i * inner + j] =
          wsub(s[i * inner + j], (T)(wadd(wadd(d[il * inner + j], d[ir * inner + j]), (T)2) >> 2));
  }
  for (size_t i = 0; i < nd; ++i) {
    size_t r = (2 * i + 2 < n) ? 2 * i + 2 : right_even(2 * i + 2, n);
    for (size_t j = 0; j < inner; ++j)
      tmp[(2 * i + 1) * inner + j] =
          wadd(d[i * inner + j], (T)(wadd(wadd(tmp[2 * i * inner + j], tmp[r * inner + j]), (T)1) >> 1));
  }
  std::memcpy(x, tmp, n * inner * sizeof(T));
}

}  // namespace

// kernel.rs:280-292
uint32_t lift_band_count(size_t n, unsigned levels) {
  uint32_t count = 0;
  size_t m = n;
  for (unsigned l = 0; l < levels; ++l) {
    if (m < 2) break;
    ++count;
    m = (m + 1) / 2;
  }
  return count;
}

namespace {

inline size_t band_len(size_t n, uint32_t k) { return (n + ((size_t)1 << k) - 1) >> k; }

template <class T> using LevelFn = void (*)(T*, size_t, size_t, T*);

// kernel.rs:307-335
template <class T> void multi_level(T* x, size_t n, size_t inner, T* tmp, unsigned levels, LevelFn<T> fn, bool fwd) {
  uint32_t bc = lift_band_count(n, levels);
  if (fwd) {
    for (uint32_t k = 0; k < bc; ++k) fn(x, band_len(n, k), inner, tmp);
  } else {
    for (uint32_t k = bc; k-- > 0;) fn(x, band_len(n, k), inner, tmp);
  }
}

size_t effective_group(uint32_t group, size_t len) { return group == 0 ? len : std::min((size_t)group, len); }

// lib.rs:156-184 (messages kept verbatim)
void check_levels(const AxisStep& s) {
  if ((s.kind == kHaar || s.kind == kLift53) && s.levels == 0) {
    std::ostringstream m;
    m << "nd_lift transform on axis \"" << s.axis << "\": kind \"" << (s.kind == kHaar ? "haar" : "lift53")
      << "\" needs levels >= 1";
    fail(kInvalidArgument, m.str());
  }
}
void check_dimension(const AxisStep& s, size_t ndim) {
  if (s.dimension >= ndim) {
    std::ostringstream m;
    m << "nd_lift transform on axis \"" << s.axis << "\": dimension " << s.dimension
      << " out of range for a " << ndim << "-dimensional chunk";
    fail(kInvalidArgument, m.str());
  }
}

// chunk.rs:83-100
void validate(size_t len, const std::vector<size_t>& shape, const std::vector<AxisStep>& steps) {
  size_t expected = 1;
  for (size_t d : shape) {
    if (d != 0 && expected > SIZE_MAX / d) fail(kInvalidArgument, "chunk shape overflows usize");
    expected *= d;
  }
  if (len != expected) fail(kInvalidArgument, "chunk length does not match the product of the shape");
  for (auto& s : steps) {
    check_dimension(s, shape.size());
    check_levels(s);
  }
}

// chunk.rs:110-150
template <class T> void apply_axis(T* chunk, size_t total, const std::vector<size_t>& shape, const AxisStep& step, bool fwd) {
  if (total == 0) return;
  size_t len = shape[step.dimension];
  if (len < 2) return;
  size_t inner = 1;
  for (size_t i = step.dimension + 1; i < shape.size(); ++i) inner *= shape[i];
  size_t group = effective_group(step.group, len);
  std::vector<T> tmp(step.kind == kDelta ? 0 : group * inner);
  for (size_t slab = 0; slab < total; slab += len * inner) {
    for (size_t g = 0; g < len; g += group) {
      size_t n = std::min(group, len - g);
      T* seg = chunk + slab + g * inner;
      switch (step.kind) {
        case kDelta: fwd ? delta_forward(seg, n, inner) : delta_inverse(seg, n, inner); break;
        case kHaar:
          multi_level<T>(seg, n, inner, tmp.data(), step.levels, fwd ? haar_forward_level<T> : haar_inverse_level<T>, fwd);
          break;
        default:
          multi_level<T>(seg, n, inner, tmp.data(), step.levels, fwd ? lift53_forward_level<T> : lift53_inverse_level<T>, fwd);
      }
    }
  }
}

struct Interval {
  i128 lo, hi;
  Interval uni(Interval o) const { return {std::min(lo, o.lo), std::max(hi, o.hi)}; }
  template <class T> bool fits() const {
    i128 mn = Plane<T>::BITS == 32 ? (i128)INT32_MIN : (i128)INT64_MIN;
    i128 mx = Plane<T>::BITS == 32 ? (i128)INT32_MAX : (i128)INT64_MAX;
    return lo >= mn && hi <= mx;
  }
};
i128 floor_div(i128 a, i128 b) {  // div_euclid for b > 0
  i128 q = a / b;
  if (a % b < 0) q -= 1;
  return q;
}

// chunk.rs:235-302
template <class T> bool step_growth(Interval input, const AxisStep& step, size_t group, Interval* out) {
  uint32_t levels = step.kind == kDelta ? 1 : lift_band_count(group, step.levels);
  Interval all = input, band = input;
  for (uint32_t l = 0; l < levels; ++l) {
    i128 lo = band.lo, hi = band.hi;
    Interval next_band, level_all;
    if (step.kind == kDelta) {
      Interval r{lo - hi, hi - lo};
      next_band = input;
      level_all = r.uni(input);
    } else if (step.kind == kHaar) {
      Interval d{lo - hi, hi - lo};
      Interval s{lo + floor_div(d.lo, 2), hi + floor_div(d.hi, 2)};
      next_band = s;
      level_all = s.uni(d);
    } else {
      Interval ps{2 * lo + 1, 2 * hi + 1};
      Interval d{lo - hi, hi - lo};
      Interval us{2 * d.lo + 2, 2 * d.hi + 2};
      if (!ps.fits<T>() || !us.fits<T>()) return false;
      Interval s{lo + floor_div(us.lo, 4), hi + floor_div(us.hi, 4)};
      next_band = s;
      level_all = s.uni(d);
    }
    all = all.uni(level_all);
    if (!all.fits<T>()) return false;
    band = next_band;
  }
  *out = all;
  return true;
}

// chunk.rs:188-231
template <class T> void check_budget(const T* chunk, size_t total, const std::vector<size_t>& shape, const std::vector<AxisStep>& steps) {
  if (total == 0) return;
  T lo = chunk[0], hi = chunk[0];
  for (size_t i = 0; i < total; ++i) { lo = std::min(lo, chunk[i]); hi = std::max(hi, chunk[i]); }
  Interval all{(i128)lo, (i128)hi};
  for (auto& s : steps) {
    size_t len = s.dimension < shape.size() ? shape[s.dimension] : 0;
    if (len < 2) continue;
    Interval grown;
    if (!step_growth<T>(all, s, effective_group(s.group, len), &grown)) {
      static const char* kn[] = {"delta", "haar", "lift53"};
      std::ostringstream m;
      m << "nd_lift overflow budget exceeded on axis \"" << s.axis << "\" (" << std::max(1u, s.levels)
        << " levels of \"" << kn[s.kind] << "\" over input range [" << i128_str(all.lo) << ", "
        << i128_str(all.hi) << "]): coefficients would leave the " << Plane<T>::BITS
        << "-bit plane; reduce levels or shorten the group";
      fail(kInvalidArgument, m.str());
    }
    all = grown;
  }
}

template <class T> void forward_t(std::vector<T>& chunk, const std::vector<size_t>& shape, const std::vector<AxisStep>& steps) {
  validate(chunk.size(), shape, steps);
  check_budget<T>(chunk.data(), chunk.size(), shape, steps);
  for (auto& s : steps) apply_axis<T>(chunk.data(), chunk.size(), shape, s, true);
}
template <class T> void inverse_t(std::vector<T>& chunk, const std::vector<size_t>& shape, const std::vector<AxisStep>& steps) {
  validate(chunk.size(), shape, steps);
  for (size_t i = steps.size(); i-- > 0;) apply_axis<T>(chunk.data(), chunk.size(), shape, steps[i], false);
}

// Strict u32 decimal parse (Rust's str::parse::<u32>).
bool parse_u32(const std::string& s, uint32_t* v) {
  if (s.empty() || s.size() > 10) return false;
  uint64_t acc = 0;
  for (char c : s) {
    if (c < '0' || c > '9') return false;
    acc = acc * 10 + uint64_t(c - '0');
  }
  if (acc > 0xFFFFFFFFull) return false;
  *v = (uint32_t)acc;
  return true;
}

template <class T> void put_le(std::vector<uint8_t>& out, T v) {
  uint8_t b[sizeof(T)];
  std::memcpy(b, &v, sizeof(T));
  out.insert(out.end(), b, b + sizeof(T));
}
template <class T> T get_le(const uint8_t* p) {
  T v;
  std::memcpy(&v, p, sizeof(T));
  return v;
}

}  // namespace

// lib.rs:228-242
void check_lift_version(const std::string& version) {
  size_t dot = version.find('.');
  std::string a = version.substr(0, dot);
  std::string b;
  if (dot != std::string::npos) {
    size_t dot2 = version.find('.', dot + 1);
    b = version.substr(dot + 1, dot2 == std::string::npos ? std::string::npos : dot2 - dot - 1);
  }
  uint32_t ma, mi;
  bool ok = parse_u32(a, &ma) && dot != std::string::npos && parse_u32(b, &mi) && ma == 0 && mi == 1;
  if (!ok)
    fail(kUnsupported, "nd_lift configuration version \"" + version +
                           "\" is not supported by this build (implements 0.1); refusing rather than mis-decoding");
}

void lift_forward_i32(std::vector<int32_t>& c, const std::vector<size_t>& s, const std::vector<AxisStep>& st) { forward_t(c, s, st); }
void lift_inverse_i32(std::vector<int32_t>& c, const std::vector<size_t>& s, const std::vector<AxisStep>& st) { inverse_t(c, s, st); }
void lift_forward_i64(std::vector<int64_t>& c, const std::vector<size_t>& s, const std::vector<AxisStep>& st) { forward_t(c, s, st); }
void lift_inverse_i64(std::vector<int64_t>& c, const std::vector<size_t>& s, const std::vector<AxisStep>& st) { inverse_t(c, s, st); }

namespace {

// lift.rs:246-294 + the PlaneConvert impls at lift.rs:137-243.
template <class P> std::vector<uint8_t> transform_bytes(const uint8_t* bytes, size_t len, const std::vector<size_t>& shape,
                                                        int dtype, const std::string& version,
                                                        const std::vector<AxisStep>& steps, bool fwd) {
  check_lift_version(version);
  for (auto& s : steps) check_levels(s);
  for (auto& s : steps) check_dimension(s, shape.size());
  size_t n = 1;
  for (size_t d : shape) {
    if (d != 0 && n > SIZE_MAX / d) fail(kInvalidArgument, "nd_lift: chunk shape overflows usize");
    n *= d;
  }
  size_t isz = dtype_size(dtype);
  if (fwd) {
    if (len != n * isz) {
      std::ostringstream m;
      m << "nd_lift encode: got " << len << " bytes for " << n << " elements of " << isz << " bytes";
      fail(kInvalidArgument, m.str());
    }
    std::vector<P> plane(n);
    for (size_t i = 0; i < n; ++i) {
      const uint8_t* p = bytes + i * isz;
      switch (dtype) {
        case kU8: plane[i] = p[0]; break;
        case kI8: plane[i] = (int8_t)p[0]; break;
        case kU16: plane[i] = get_le<uint16_t>(p); break;
        case kI16: plane[i] = get_le<int16_t>(p); break;
        case kI32: plane[i] = get_le<int32_t>(p); break;
        case kI64: plane[i] = (P)get_le<int64_t>(p); break;
        case kU32: case kU64: break;
      }
    }
    if (dtype == kU32 || dtype == kU64) {
      uint64_t mx = 0;
      for (size_t i = 0; i < n; ++i) {
        uint64_t v = dtype == kU32 ? get_le<uint32_t>(bytes + i * 4) : get_le<uint64_t>(bytes + i * 8);
        mx = std::max(mx, v);
      }
      uint64_t pmax = dtype == kU32 ? (uint64_t)INT32_MAX : (uint64_t)INT64_MAX;
      if (mx > pmax) {
        std::ostringstream m;
        m << "nd_lift overflow budget: input value " << mx << " does not fit the widened "
          << (dtype == kU32 ? "i32" : "i64") << " coefficient plane";
        fail(kInvalidArgument, m.str());
      }
      for (size_t i = 0; i < n; ++i)
        plane[i] = (P)(dtype == kU32 ? get_le<uint32_t>(bytes + i * 4) : get_le<uint64_t>(bytes + i * 8));
    }
    forward_t(plane, shape, steps);
    std::vector<uint8_t> out;
    out.reserve(n * sizeof(P));
    for (P v : plane) put_le(out, v);
    return out;
  }
  if (len != n * sizeof(P)) {
    std::ostringstream m;
    m << "nd_lift decode: got " << len << " bytes for " << n << " coefficients of " << sizeof(P) << " bytes";
    fail(kInvalidArgument, m.str());
  }
  std::vector<P> plane(n);
  for (size_t i = 0; i < n; ++i) plane[i] = get_le<P>(bytes + i * sizeof(P));
  inverse_t(plane, shape, steps);
  std::vector<uint8_t> out;
  out.reserve(n * isz);
  if (dtype == kI32 || dtype == kI64) {
    for (P v : plane) put_le(out, v);
    return out;
  }
  i128 lo_b, hi_b;
  switch (dtype) {
    case kU8: lo_b = 0; hi_b = 255; break;
    case kI8: lo_b = -128; hi_b = 127; break;
    case kU16: lo_b = 0; hi_b = 65535; break;
    case kI16: lo_b = -32768; hi_b = 32767; break;
    default: lo_b = 0; hi_b = dtype == kU32 ? (i128)INT32_MAX : (i128)INT64_MAX; break;
  }
  P lo = 0, hi = 0;
  if (!plane.empty()) {
    lo = *std::min_element(plane.begin(), plane.end());
    hi = *std::max_element(plane.begin(), plane.end());
  }
  if (!plane.empty() && ((i128)lo < lo_b || (i128)hi > hi_b)) {
    std::ostringstream m;
    m << "nd_lift decode: coefficient range [" << (long long)lo << ", " << (long long)hi
      << "] does not narrow back to " << dtype_name(dtype) << " (corrupt or mismatched chunk)";
    fail(kInvalidArgument, m.str());
  }
  for (P v : plane) {
    switch (dtype) {
      case kU8: case kI8: out.push_back((uint8_t)v); break;
      case kU16: case kI16: put_le<uint16_t>(out, (uint16_t)v); break;
      case kU32: put_le<uint32_t>(out, (uint32_t)v); break;
      case kU64: put_le<uint64_t>(out, (uint64_t)v); break;
    }
  }
  return out;
}

}  // namespace

std::vector<uint8_t> lift_forward_chunk(const uint8_t* b, size_t len, const std::vector<size_t>& shape, int dtype,
                                        const std::string& version, const std::vector<AxisStep>& steps) {
  if (dtype == kU64 || dtype == kI64) return transform_bytes<int64_t>(b, len, shape, dtype, version, steps, true);
  if (dtype == kF32 || dtype == kF64) fail(kUnsupported, "nd_lift: floating-point dtypes are not supported");
  return transform_bytes<int32_t>(b, len, shape, dtype, version, steps, true);
}
std::vector<uint8_t> lift_inverse_chunk(const uint8_t* b, size_t len, const std::vector<size_t>& shape, int dtype,
                                        const std::string& version, const std::vector<AxisStep>& steps) {
  if (dtype == kU64 || dtype == kI64) return transform_bytes<int64_t>(b, len, shape, dtype, version, steps, false);
  if (dtype == kF32 || dtype == kF64) fail(kUnsupported, "nd_lift: floating-point dtypes are not supported");
  return transform_bytes<int32_t>(b, len, shape, dtype, version, steps, false);
}

}  // namespace orc
