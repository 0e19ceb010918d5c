// Tape wire format: the serde JSON form of BytecodeTape
// (src/bytecode_tape/serde_support.rs:11-100), so that tapes recorded and
// serialized by the Rust crate load into this library's host code, and
// tapes recorded here travel back.
//
// Writer: serde_json's compact layout of the eight fields in declaration
// order -- opcodes as variant-name strings, arg_indices as [a, b] pairs,
// values as shortest round-trip decimals in ryu's layout ("1.0", "0.25",
// "1e20", "-0.0"), non-finite values as `null` (serde_json's rule), and the
// empty custom_second_args map.  A tape holding a Custom opcode is refused,
// as serde_support.rs:13-18 refuses it.
//
// Reader: a strict JSON parser driven by the TapeData schema
// (serde_support.rs:36-47): unknown fields are skipped, output_indices and
// custom_second_args default to empty, a missing required field or a
// duplicated one is an error; then the deserializer's own checks in its
// order (Custom opcode, custom_second_args, num_variables stamp,
// serde_support.rs:55-80) and BytecodeTape::validate
// (src/bytecode_tape/mod.rs:575-703) with the reference's messages.
// Decimal values are converted with correct rounding; an f32 tape's values
// are parsed as f64 and then rounded to f32, as serde_json does for f32.
//
// The same data model is also read and written as CBOR (RFC 8949), the
// binary form the reference round-trips through ciborium
// (tests/serde.rs:156-170), with the same checks and messages.
#include <clocale>
#include <limits>
#include <cmath>
#include <cstdlib>
#include <locale.h>

#include "ehip_internal.h"

namespace {

// OpCode variant names by discriminant (src/opcode.rs:24-130)
const char* const kOpName[] = {
    "Input", "Const", "Add",   "Sub",   "Mul",   "Div",   "Rem",   "Powf",  "Atan2",  "Hypot", "Max",
    "Min",   "Neg",   "Recip", "Sqrt",  "Cbrt",  "Powi",  "Exp",   "Exp2",  "ExpM1",  "Ln",    "Log2",
    "Log10", "Ln1p",  "Sin",   "Cos",   "Tan",   "Asin",  "Acos",  "Atan",  "Sinh",   "Cosh",  "Tanh",
    "Asinh", "Acosh", "Atanh", "Abs",   "Signum", "Floor", "Ceil", "Round", "Trunc",  "Fract", "Custom"};
constexpr uint32_t kNumOps = sizeof(kOpName) / sizeof(kOpName[0]);
constexpr uint32_t OP_INPUT_ = 0, OP_CONST_ = 1, OP_POWI_ = 16, OP_CUSTOM_ = 43;
constexpr uint32_t UNUSED = 0xFFFFFFFFu;  // empty operand slot (src/bytecode_tape/mod.rs)

bool binary_op(uint32_t op) { return op >= 2 && op <= 11; }  // Add..Min

// ---- number formatting ------------------------------------------------------

// Shortest round-trip digits and decimal exponent of a finite positive
// value: v = digits * 10^exp10, digits without trailing zeros.  For each
// length L the correctly rounded L-digit decimal is tried, then its two
// neighbours: at a power of two the rounding interval is wider above v than
// below, so the nearest L-digit decimal can miss it while the next one up
// lies inside.  The first length with a hit gives the shortest string; the
// nearest hit is kept.
template <typename F>
void shortest(F v, std::string& digits, int& exp10) {
  char buf[64];
  const int maxp = sizeof(F) == 8 ? 17 : 9;
  auto reads_back = [&](unsigned long long D, int ex) {
    char t[64];
    std::snprintf(t, sizeof t, "%llue%d", D, ex);
    F back = sizeof(F) == 8 ? (F)std::strtod(t, nullptr) : (F)std::strtof(t, nullptr);
    return back == v;
  };
  unsigned long long best = 0;
  int best_ex = 0;
  for (int p = 0; p < maxp; ++p) {
    std::snprintf(buf, sizeof buf, "%.*e", p, (double)v);
    unsigned long long D = 0;
    const char* q = buf;
    for (; *q && *q != 'e'; ++q)
      if (*q != '.') D = D * 10 + (unsigned long long)(*q - '0');
    const int ex = std::atoi(q + 1) - p;  // v ~ D * 10^ex
    if (reads_back(D, ex)) {
      best = D, best_ex = ex;
      break;
    }
    if (reads_back(D + 1, ex)) {
      best = D + 1, best_ex = ex;
      break;
    }
    if (D > 1 && reads_back(D - 1, ex)) {
      best = D - 1, best_ex = ex;
      break;
    }
    best = D, best_ex = ex;  // the last length always reads back
  }
  while (best % 10 == 0 && best) {
    best /= 10;
    ++best_ex;
  }
  digits = std::to_string(best);
  exp10 = best_ex;
}

// ryu's pretty layout (the formatter serde_json uses): fixed notation for
// decimal exponents in a window around 1, scientific outside it.
template <typename F>
void write_float(std::string& out, F v) {
  if (!std::isfinite(v)) {
    out += "null";
    return;
  }
  if (std::signbit(v)) out.push_back('-');
  if (v == 0) {
    out += "0.0";
    return;
  }
  std::string d;
  int k;
  shortest<F>(std::fabs(v), d, k);
  const int len = (int)d.size();
  const int kk = len + k;  // 10^(kk-1) <= |v| < 10^kk
  const int hi = sizeof(F) == 8 ? 16 : 13, lo = sizeof(F) == 8 ? -5 : -6;
  if (k >= 0 && kk <= hi) {
    out += d;
    out.append((size_t)k, '0');
    out += ".0";
  } else if (kk > 0 && kk <= hi) {
    out.append(d, 0, (size_t)kk);
    out.push_back('.');
    out.append(d, (size_t)kk, std::string::npos);
  } else if (kk > lo && kk <= 0) {
    out += "0.";
    out.append((size_t)(-kk), '0');
    out += d;
  } else {
    out.push_back(d[0]);
    if (len > 1) {
      out.push_back('.');
      out.append(d, 1, std::string::npos);
    }
    out.push_back('e');
    out += std::to_string(kk - 1);
  }
}

// ---- reader -----------------------------------------------------------------

struct Reader {
  const char* b;
  const char* p;
  const char* e;
  std::string err;

  bool fail(const char* fmt, ...) {
    char buf[400];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    char at[48];
    std::snprintf(at, sizeof at, " at byte %zu", (size_t)(p - b));
    err = std::string(buf) + at;
    return false;
  }
  void ws() {
    while (p < e && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
  }
  bool eat(char c) {
    ws();
    if (p < e && *p == c) {
      ++p;
      return true;
    }
    return false;
  }
  bool expect(char c, const char* what) {
    if (eat(c)) return true;
    if (p >= e) return fail("EOF while parsing %s", what);
    return fail("expected `%c` while parsing %s", c, what);
  }
  const char* kind() {
    ws();
    if (p >= e) return "EOF";
    switch (*p) {
      case 'n': return "null";
      case 't': case 'f': return "boolean";
      case '"': return "string";
      case '[': return "sequence";
      case '{': return "map";
      default: return "number";
    }
  }

  static void put_utf8(std::string& s, uint32_t cp) {
    if (cp < 0x80) {
      s.push_back((char)cp);
    } else if (cp < 0x800) {
      s.push_back((char)(0xC0 | (cp >> 6)));
      s.push_back((char)(0x80 | (cp & 0x3F)));
    } else if (cp < 0x10000) {
      s.push_back((char)(0xE0 | (cp >> 12)));
      s.push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
      s.push_back((char)(0x80 | (cp & 0x3F)));
    } else {
      s.push_back((char)(0xF0 | (cp >> 18)));
      s.push_back((char)(0x80 | ((cp >> 12) & 0x3F)));
      s.push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
      s.push_back((char)(0x80 | (cp & 0x3F)));
    }
  }
  bool hex4(uint32_t& v) {
    if (e - p < 4) return fail("EOF while parsing a string escape");
    v = 0;
    for (int i = 0; i < 4; ++i) {
      char c = *p++;
      v <<= 4;
      if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
      else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
      else return fail("invalid escape");
    }
    return true;
  }
  // A JSON string (the opening quote already checked by the caller's ws()).
  bool string(std::string& s) {
    ws();
    if (p >= e || *p != '"') return fail("invalid type: %s, expected a string", kind());
    ++p;
    s.clear();
    while (true) {
      if (p >= e) return fail("EOF while parsing a string");
      unsigned char c = (unsigned char)*p++;
      if (c == '"') return true;
      if (c < 0x20) return fail("control character (\\u0000-\\u001F) found while parsing a string");
      if (c != '\\') {
        s.push_back((char)c);
        continue;
      }
      if (p >= e) return fail("EOF while parsing a string");
      char x = *p++;
      switch (x) {
        case '"': s.push_back('"'); break;
        case '\\': s.push_back('\\'); break;
        case '/': s.push_back('/'); break;
        case 'b': s.push_back('\b'); break;
        case 'f': s.push_back('\f'); break;
        case 'n': s.push_back('\n'); break;
        case 'r': s.push_back('\r'); break;
        case 't': s.push_back('\t'); break;
        case 'u': {
          uint32_t cp;
          if (!hex4(cp)) return false;
          if (cp >= 0xD800 && cp < 0xDC00) {
            uint32_t lo;
            if (e - p < 2 || p[0] != '\\' || p[1] != 'u') return fail("lone leading surrogate in hex escape");
            p += 2;
            if (!hex4(lo)) return false;
            if (lo < 0xDC00 || lo >= 0xE000) return fail("lone leading surrogate in hex escape");
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          } else if (cp >= 0xDC00 && cp < 0xE000) {
            return fail("lone trailing surrogate in hex escape");
          }
          put_utf8(s, cp);
          break;
        }
        default: return fail("invalid escape");
      }
    }
  }

  // One JSON number token; `integer` is false when it has a fraction or an
  // exponent.  Returns the token's extent.
  bool number(const char*& t0, const char*& t1, bool& integer, bool& negative) {
    ws();
    t0 = p;
    negative = false;
    if (p < e && *p == '-') {
      negative = true;
      ++p;
    }
    if (p >= e || !(*p >= '0' && *p <= '9')) {
      p = t0;
      return fail("invalid type: %s, expected a number", kind());
    }
    if (*p == '0') {
      ++p;
      if (p < e && *p >= '0' && *p <= '9') return fail("invalid number");
    } else {
      while (p < e && *p >= '0' && *p <= '9') ++p;
    }
    integer = true;
    if (p < e && *p == '.') {
      integer = false;
      ++p;
      if (p >= e || !(*p >= '0' && *p <= '9')) return fail("invalid number");
      while (p < e && *p >= '0' && *p <= '9') ++p;
    }
    if (p < e && (*p == 'e' || *p == 'E')) {
      integer = false;
      ++p;
      if (p < e && (*p == '+' || *p == '-')) ++p;
      if (p >= e || !(*p >= '0' && *p <= '9')) return fail("invalid number");
      while (p < e && *p >= '0' && *p <= '9') ++p;
    }
    t1 = p;
    return true;
  }

  bool u32(uint32_t& v, const char* field) {
    ws();
    if (p < e && !(*p == '-' || (*p >= '0' && *p <= '9')))
      return fail("invalid type: %s, expected u32 (field `%s`)", kind(), field);
    const char *t0, *t1;
    bool integer, neg;
    if (!number(t0, t1, integer, neg)) return false;
    std::string tok(t0, t1);
    if (!integer || (t1 - t0) > 20 + (neg ? 1 : 0))
      return fail("invalid type: floating point `%s`, expected u32 (field `%s`)", tok.c_str(), field);
    if (neg) {
      bool zero = tok == "-0";
      if (!zero) return fail("invalid value: integer `%s`, expected u32 (field `%s`)", tok.c_str(), field);
      v = 0;
      return true;
    }
    unsigned long long x = std::strtoull(tok.c_str(), nullptr, 10);
    if (x > 0xFFFFFFFFull || tok.size() > 10)
      return fail("invalid value: integer `%s`, expected u32 (field `%s`)", tok.c_str(), field);
    v = (uint32_t)x;
    return true;
  }

  bool f64(double& v) {
    ws();
    if (p < e && !(*p == '-' || (*p >= '0' && *p <= '9'))) return fail("invalid type: %s, expected f64", kind());
    const char *t0, *t1;
    bool integer, neg;
    if (!number(t0, t1, integer, neg)) return false;
    static locale_t cloc = newlocale(LC_NUMERIC_MASK, "C", (locale_t)0);
    char small[64];
    std::string big;
    const char* s;
    size_t n = (size_t)(t1 - t0);
    if (n < sizeof small) {
      std::memcpy(small, t0, n);
      small[n] = 0;
      s = small;
    } else {
      big.assign(t0, t1);
      s = big.c_str();
    }
    v = strtod_l(s, nullptr, cloc);
    if (std::isinf(v)) return fail("number out of range");
    return true;
  }

  bool skip(int depth) {
    if (depth > 128) return fail("recursion limit exceeded");
    ws();
    if (p >= e) return fail("EOF while parsing a value");
    char c = *p;
    if (c == '"') {
      std::string s;
      return string(s);
    }
    if (c == '{') {
      ++p;
      if (eat('}')) return true;
      do {
        std::string k;
        if (!string(k) || !expect(':', "an object") || !skip(depth + 1)) return false;
      } while (eat(','));
      return expect('}', "an object");
    }
    if (c == '[') {
      ++p;
      if (eat(']')) return true;
      do {
        if (!skip(depth + 1)) return false;
      } while (eat(','));
      return expect(']', "a list");
    }
    auto lit = [&](const char* w) {
      size_t n = std::strlen(w);
      if ((size_t)(e - p) >= n && std::memcmp(p, w, n) == 0) {
        p += n;
        return true;
      }
      return fail("expected value");
    };
    if (c == 't') return lit("true");
    if (c == 'f') return lit("false");
    if (c == 'n') return lit("null");
    const char *t0, *t1;
    bool integer, neg;
    return number(t0, t1, integer, neg);
  }

  // An OpCode: "Name", or the map form {"Name": null} serde_json also
  // accepts for a unit variant.
  bool opcode(uint32_t& op) {
    ws();
    std::string name;
    bool map = false;
    if (p < e && *p == '{') {
      ++p;
      map = true;
    } else if (p < e && *p != '"') {
      return fail("invalid type: %s, expected an OpCode variant name", kind());
    }
    if (!string(name)) return false;
    op = kNumOps;
    for (uint32_t i = 0; i < kNumOps; ++i)
      if (name == kOpName[i]) op = i;
    if (op == kNumOps) return fail("unknown variant `%s`", name.c_str());
    if (map) {
      if (!expect(':', "an enum")) return false;
      ws();
      if (!((size_t)(e - p) >= 4 && std::memcmp(p, "null", 4) == 0))
        return fail("invalid type: %s, expected unit variant OpCode::%s", kind(), name.c_str());
      p += 4;
      if (!expect('}', "an enum")) return false;
    }
    return true;
  }

  template <typename Fn>
  bool array(Fn&& elem, const char* field) {
    ws();
    if (p >= e || *p != '[') return fail("invalid type: %s, expected a sequence (field `%s`)", kind(), field);
    ++p;
    if (eat(']')) return true;
    do {
      if (!elem()) return false;
    } while (eat(','));
    return expect(']', "a list");
  }
};

struct TapeData {
  std::vector<uint32_t> opcodes, arg0, arg1, output_indices;
  std::vector<double> values;
  uint32_t num_inputs = 0, num_variables = 0, output_index = 0;
  size_t n_custom_second = 0;
  uint32_t first_custom_key = 0, first_custom_val = 0;
};

// TapeData deserialization (serde_support.rs:36-49).
bool parse_tape(Reader& r, TapeData& d) {
  enum { F_OPC, F_ARGS, F_VALS, F_NIN, F_NVAR, F_OUT, F_OUTS, F_CSA, NF };
  static const char* const names[NF] = {"opcodes",      "arg_indices",   "values",         "num_inputs",
                                        "num_variables", "output_index", "output_indices", "custom_second_args"};
  bool seen[NF] = {};
  r.ws();
  if (r.p >= r.e || *r.p != '{') return r.fail("invalid type: %s, expected struct TapeData", r.kind());
  ++r.p;
  if (!r.eat('}')) {
    do {
      std::string key;
      if (!r.string(key) || !r.expect(':', "an object")) return false;
      int f = -1;
      for (int i = 0; i < NF; ++i)
        if (key == names[i]) f = i;
      if (f < 0) {
        if (!r.skip(1)) return false;
        continue;
      }
      if (seen[f]) return r.fail("duplicate field `%s`", names[f]);
      seen[f] = true;
      bool ok = true;
      switch (f) {
        case F_OPC:
          ok = r.array([&] {
            uint32_t op;
            if (!r.opcode(op)) return false;
            d.opcodes.push_back(op);
            return true;
          }, "opcodes");
          break;
        case F_ARGS:
          ok = r.array([&] {
            uint32_t a = 0, b = 0;
            r.ws();
            if (r.p >= r.e || *r.p != '[')
              return r.fail("invalid type: %s, expected an array of length 2", r.kind());
            ++r.p;
            if (r.eat(']')) return r.fail("invalid length 0, expected an array of length 2");
            if (!r.u32(a, "arg_indices")) return false;
            if (!r.eat(',')) return r.fail("invalid length 1, expected an array of length 2");
            if (!r.u32(b, "arg_indices")) return false;
            if (!r.eat(']')) return r.fail("trailing characters: arg_indices entry longer than 2");
            d.arg0.push_back(a);
            d.arg1.push_back(b);
            return true;
          }, "arg_indices");
          break;
        case F_VALS:
          ok = r.array([&] {
            double v;
            if (!r.f64(v)) return false;
            d.values.push_back(v);
            return true;
          }, "values");
          break;
        case F_NIN: ok = r.u32(d.num_inputs, "num_inputs"); break;
        case F_NVAR: ok = r.u32(d.num_variables, "num_variables"); break;
        case F_OUT: ok = r.u32(d.output_index, "output_index"); break;
        case F_OUTS:
          ok = r.array([&] {
            uint32_t v;
            if (!r.u32(v, "output_indices")) return false;
            d.output_indices.push_back(v);
            return true;
          }, "output_indices");
          break;
        case F_CSA: {
          r.ws();
          if (r.p >= r.e || *r.p != '{')
            return r.fail("invalid type: %s, expected a map (field `custom_second_args`)", r.kind());
          ++r.p;
          if (r.eat('}')) break;
          do {
            std::string k;
            uint32_t v;
            if (!r.string(k)) return false;
            char* end = nullptr;
            unsigned long long kv = std::strtoull(k.c_str(), &end, 10);
            if (k.empty() || *end || k[0] == '-' || k[0] == '+' || kv > 0xFFFFFFFFull)
              return r.fail("invalid value: map key `%s`, expected u32", k.c_str());
            if (!r.expect(':', "an object") || !r.u32(v, "custom_second_args")) return false;
            if (d.n_custom_second++ == 0) {
              d.first_custom_key = (uint32_t)kv;
              d.first_custom_val = v;
            }
          } while (r.eat(','));
          ok = r.expect('}', "an object");
          break;
        }
      }
      if (!ok) return false;
    } while (r.eat(','));
    if (!r.expect('}', "an object")) return false;
  }
  for (int f = 0; f <= F_OUT; ++f)
    if (!seen[f]) return r.fail("missing field `%s`", names[f]);
  r.ws();
  if (r.p != r.e) return r.fail("trailing characters");
  return true;
}

std::string pair_str(uint32_t a, uint32_t b) { return "[" + std::to_string(a) + ", " + std::to_string(b) + "]"; }

// Post-parse checks of the deserializer (serde_support.rs:55-80) and
// BytecodeTape::validate (src/bytecode_tape/mod.rs:575-703).  The tape has
// no custom-op registry after deserialization, so a Custom entry can only
// appear here if the first check were skipped.
std::string check_tape(const TapeData& d) {
  char buf[400];
  auto err = [&](const char* fmt, auto... xs) {
    std::snprintf(buf, sizeof buf, fmt, xs...);
    return std::string(buf);
  };
  const size_t nv = d.opcodes.size();
  for (size_t i = 0; i < nv; ++i)
    if (d.opcodes[i] == OP_CUSTOM_)
      return err("opcodes[%zu] is Custom, which cannot be deserialized (custom ops have no serializable callback)",
                 i);
  if (d.n_custom_second) return "custom_second_args must be empty: custom ops cannot be deserialized";
  if ((size_t)d.num_variables != nv) return err("num_variables (%u) != opcodes.len() (%zu)", d.num_variables, nv);
  // validate()
  if (d.arg0.size() != nv) return err("arg_indices.len() (%zu) != opcodes.len() (%zu)", d.arg0.size(), nv);
  if (d.values.size() != nv) return err("values.len() (%zu) != opcodes.len() (%zu)", d.values.size(), nv);
  const size_t ni = d.num_inputs;
  if (ni > nv) return err("num_inputs (%zu) > tape length (%zu)", ni, nv);
  for (size_t i = 0; i < nv; ++i) {
    uint32_t op = d.opcodes[i];
    if ((op == OP_INPUT_) != (i < ni)) {
      if (i < ni) return err("opcodes[%zu] should be Input but is %s", i, kOpName[op]);
      return err("opcodes[%zu] is Input, but only the first %zu entries may be inputs", i, ni);
    }
    if (i < ni && (d.arg0[i] != UNUSED || d.arg1[i] != UNUSED))
      return err("arg_indices[%zu] must be [UNUSED, UNUSED] for an Input entry, found %s", i,
                 pair_str(d.arg0[i], d.arg1[i]).c_str());
  }
  if ((size_t)d.output_index >= nv) return err("output_index (%u) >= tape length (%zu)", d.output_index, nv);
  for (size_t j = 0; j < d.output_indices.size(); ++j)
    if ((size_t)d.output_indices[j] >= nv)
      return err("output_indices[%zu] (%u) >= tape length (%zu)", j, d.output_indices[j], nv);
  for (size_t i = ni; i < nv; ++i) {
    uint32_t op = d.opcodes[i];
    uint32_t a = d.arg0[i], b = d.arg1[i];
    if (op == OP_CONST_) {
      if (a != UNUSED || b != UNUSED)
        return err("arg_indices[%zu] must be [UNUSED, UNUSED] for a Const entry, found %s", i,
                   pair_str(a, b).c_str());
      continue;
    }
    if ((size_t)a >= i) return err("arg_indices[%zu][0] (%u) must reference an earlier slot", i, a);
    if (op == OP_POWI_) continue;  // slot 1 holds the exponent's bits
    if (binary_op(op)) {
      if ((size_t)b >= i) return err("arg_indices[%zu][1] (%u) must reference an earlier slot", i, b);
    } else if (b != UNUSED) {
      return err("arg_indices[%zu][1] (%u) should be UNUSED for unary op %s", i, b, kOpName[op]);
    }
  }
  return "";
}

int32_t read_tape(const char* json, uint64_t len, TapeData& d) {
  if (!json && len) return fail(EHIP_OTHER, "tape JSON: null buffer");
  Reader r{json, json, json + len, {}};
  if (!parse_tape(r, d)) return fail(EHIP_OTHER, "tape JSON: %s", r.err.c_str());
  std::string msg = check_tape(d);
  if (!msg.empty()) return fail(EHIP_OTHER, "tape JSON: %s", msg.c_str());
  return EHIP_OK;
}

// ---- CBOR (ciborium's layout of the same serde data model) ------------------
//
// RFC 8949 encoding as ciborium writes it: the struct as a definite map
// with text keys in field order, unit variants as text strings, sequences
// as definite arrays, unsigned integers in their shortest head, and floats
// in the shortest of binary16 / binary32 / binary64 that holds the value
// exactly (NaN as the binary16 quiet NaN 0x7e00).  Unlike JSON, CBOR
// carries Inf and NaN, so such tapes round-trip.

void cbor_head(std::string& o, uint32_t major, uint64_t v) {
  const char m = (char)(major << 5);
  if (v < 24) {
    o.push_back((char)(m | (char)v));
  } else if (v <= 0xFF) {
    o.push_back((char)(m | 24));
    o.push_back((char)v);
  } else if (v <= 0xFFFF) {
    o.push_back((char)(m | 25));
    for (int s = 8; s >= 0; s -= 8) o.push_back((char)(v >> s));
  } else if (v <= 0xFFFFFFFFull) {
    o.push_back((char)(m | 26));
    for (int s = 24; s >= 0; s -= 8) o.push_back((char)(v >> s));
  } else {
    o.push_back((char)(m | 27));
    for (int s = 56; s >= 0; s -= 8) o.push_back((char)(v >> s));
  }
}

void cbor_text(std::string& o, const char* t) {
  const size_t n = std::strlen(t);
  cbor_head(o, 3, n);
  o.append(t, n);
}

// binary16 bits of v when v is exactly representable there.
bool half_of(double v, uint16_t& h) {
  const uint16_t sign = std::signbit(v) ? 0x8000 : 0;
  const double a = std::fabs(v);
  if (a == 0) { h = sign; return true; }
  if (std::isinf(a)) { h = sign | 0x7C00; return true; }
  int e;
  const double fr = std::frexp(a, &e);  // a = fr * 2^e, fr in [0.5, 1)
  const int ex = e - 1;                 // a = (2 fr) * 2^ex
  if (ex >= -14 && ex <= 15) {          // normal: 10 fraction bits
    const double mant = (2 * fr - 1) * 1024;
    if (mant != std::floor(mant)) return false;
    h = sign | (uint16_t)((ex + 15) << 10) | (uint16_t)mant;
    return true;
  }
  if (ex < -14 && ex >= -24) {  // subnormal: a = k * 2^-24
    const double k = std::ldexp(a, 24);
    if (k != std::floor(k)) return false;
    h = sign | (uint16_t)k;
    return true;
  }
  return false;
}

double half_value(uint16_t h) {
  const int e = (h >> 10) & 0x1F, f = h & 0x3FF;
  double v;
  if (e == 0) v = std::ldexp((double)f, -24);
  else if (e == 31) v = f ? std::numeric_limits<double>::quiet_NaN() : std::numeric_limits<double>::infinity();
  else v = std::ldexp(1.0 + f / 1024.0, e - 15);
  return (h & 0x8000) ? -v : v;
}

void cbor_float(std::string& o, double v) {
  uint16_t h;
  if (std::isnan(v)) {
    o.push_back((char)0xF9);
    o.push_back((char)0x7E);
    o.push_back((char)0x00);
  } else if (half_of(v, h)) {
    o.push_back((char)0xF9);
    o.push_back((char)(h >> 8));
    o.push_back((char)h);
  } else if ((double)(float)v == v) {
    const float f = (float)v;
    uint32_t b;
    std::memcpy(&b, &f, 4);
    o.push_back((char)0xFA);
    for (int s = 24; s >= 0; s -= 8) o.push_back((char)(b >> s));
  } else {
    uint64_t b;
    std::memcpy(&b, &v, 8);
    o.push_back((char)0xFB);
    for (int s = 56; s >= 0; s -= 8) o.push_back((char)(b >> s));
  }
}

struct CborReader {
  const uint8_t* b;
  const uint8_t* p;
  const uint8_t* e;
  std::string err;

  bool fail(const char* fmt, ...) {
    char buf[300];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    char at[48];
    std::snprintf(at, sizeof at, " at byte %zu", (size_t)(p - b));
    err = std::string(buf) + at;
    return false;
  }
  static const char* kind(uint32_t major) {
    static const char* const k[8] = {"integer", "negative integer", "bytes", "string", "array", "map", "tag",
                                     "float or simple value"};
    return k[major & 7];
  }
  // Head of the next item: major type, argument, and whether the length is
  // indefinite (additional information 31).
  bool head(uint32_t& major, uint64_t& arg, bool& indef) {
    if (p >= e) return fail("unexpected end of input");
    const uint8_t ib = *p++;
    major = ib >> 5;
    const uint32_t ai = ib & 31;
    indef = false;
    if (ai < 24) { arg = ai; return true; }
    if (ai == 31) {
      if (major == 0 || major == 1 || major == 6) return fail("invalid indefinite length");
      indef = true;
      arg = 0;
      return true;
    }
    if (ai > 27) return fail("reserved additional information %u", ai);
    const int n = 1 << (ai - 24);
    if (e - p < n) return fail("unexpected end of input");
    arg = 0;
    for (int i = 0; i < n; ++i) arg = (arg << 8) | *p++;
    return true;
  }
  bool peek_break() { return p < e && *p == 0xFF; }
  bool u32(uint32_t& v, const char* field) {
    uint32_t mj; uint64_t a; bool ind;
    if (!head(mj, a, ind)) return false;
    if (mj == 1) return fail("invalid value: negative integer, expected u32 (field `%s`)", field);
    if (mj != 0) return fail("invalid type: %s, expected u32 (field `%s`)", kind(mj), field);
    if (a > 0xFFFFFFFFull) return fail("invalid value: integer `%llu`, expected u32 (field `%s`)",
                                       (unsigned long long)a, field);
    v = (uint32_t)a;
    return true;
  }
  bool f64(double& v) {
    const uint8_t* at = p;
    uint32_t mj; uint64_t a; bool ind;
    if (!head(mj, a, ind)) return false;
    if (mj == 0) { v = (double)a; return true; }
    if (mj == 1) { v = -1.0 - (double)a; return true; }
    if (mj == 7) {
      const uint8_t ai = *at & 31;
      if (ai == 25) { v = half_value((uint16_t)a); return true; }
      if (ai == 26) { uint32_t b = (uint32_t)a; float f; std::memcpy(&f, &b, 4); v = f; return true; }
      if (ai == 27) { std::memcpy(&v, &a, 8); return true; }
    }
    return fail("invalid type: %s, expected f64", kind(mj));
  }
  bool text(std::string& s) {
    uint32_t mj; uint64_t a; bool ind;
    if (!head(mj, a, ind)) return false;
    if (mj != 3) return fail("invalid type: %s, expected a string", kind(mj));
    if (ind) return fail("indefinite-length strings are not supported");
    if ((uint64_t)(e - p) < a) return fail("unexpected end of input");
    s.assign((const char*)p, (size_t)a);
    p += a;
    return true;
  }
  bool skip(int depth) {
    if (depth > 128) return fail("recursion limit exceeded");
    uint32_t mj; uint64_t a; bool ind;
    if (!head(mj, a, ind)) return false;
    switch (mj) {
      case 0: case 1: return true;
      case 2: case 3:
        if (ind) {
          while (!peek_break()) if (!skip(depth + 1)) return false;
          ++p;
          return true;
        }
        if ((uint64_t)(e - p) < a) return fail("unexpected end of input");
        p += a;
        return true;
      case 4: case 5: {
        const uint64_t per = mj == 5 ? 2 : 1;
        if (ind) {
          while (!peek_break()) if (!skip(depth + 1)) return false;
          ++p;
          return true;
        }
        for (uint64_t i = 0; i < a * per; ++i) if (!skip(depth + 1)) return false;
        return true;
      }
      case 6: return skip(depth + 1);
      default: return true;  // simple values and floats: head only
    }
  }
  // Items of an array (definite or indefinite), each read by elem().
  template <typename Fn>
  bool array(Fn&& elem, const char* field) {
    uint32_t mj; uint64_t a; bool ind;
    if (!head(mj, a, ind)) return false;
    if (mj != 4) return fail("invalid type: %s, expected a sequence (field `%s`)", kind(mj), field);
    if (ind) {
      while (!peek_break()) if (!elem()) return false;
      ++p;
      return true;
    }
    // every item takes at least one byte
    if (a > (uint64_t)(e - p))
      return fail("unexpected end of input: %llu items declared for field `%s`", (unsigned long long)a, field);
    for (uint64_t i = 0; i < a; ++i) if (!elem()) return false;
    return true;
  }
  bool opcode(uint32_t& op) {
    std::string name;
    if (p < e && (*p >> 5) == 5) {  // {"Name": null}
      uint32_t mj; uint64_t a; bool ind;
      if (!head(mj, a, ind)) return false;
      if (ind || a != 1) return fail("invalid length, expected a one-entry map for an enum");
      if (!text(name)) return false;
      if (p >= e || *p != 0xF6) return fail("invalid type, expected unit variant OpCode::%s", name.c_str());
      ++p;
    } else if (!text(name)) {
      return false;
    }
    op = kNumOps;
    for (uint32_t i = 0; i < kNumOps; ++i)
      if (name == kOpName[i]) op = i;
    if (op == kNumOps) return fail("unknown variant `%s`", name.c_str());
    return true;
  }
};

bool parse_tape_cbor(CborReader& r, TapeData& d) {
  enum { F_OPC, F_ARGS, F_VALS, F_NIN, F_NVAR, F_OUT, F_OUTS, F_CSA, NF };
  static const char* const names[NF] = {"opcodes",      "arg_indices",   "values",         "num_inputs",
                                        "num_variables", "output_index", "output_indices", "custom_second_args"};
  bool seen[NF] = {};
  uint32_t mj; uint64_t n; bool ind;
  if (!r.head(mj, n, ind)) return false;
  if (mj != 5) return r.fail("invalid type: %s, expected struct TapeData", CborReader::kind(mj));
  for (uint64_t k = 0; ind || k < n; ++k) {
    if (ind && r.peek_break()) { ++r.p; break; }
    std::string key;
    if (!r.text(key)) return false;
    int f = -1;
    for (int i = 0; i < NF; ++i)
      if (key == names[i]) f = i;
    if (f < 0) {
      if (!r.skip(1)) return false;
      continue;
    }
    if (seen[f]) return r.fail("duplicate field `%s`", names[f]);
    seen[f] = true;
    bool ok = true;
    switch (f) {
      case F_OPC:
        ok = r.array([&] { uint32_t op; if (!r.opcode(op)) return false; d.opcodes.push_back(op); return true; },
                     "opcodes");
        break;
      case F_ARGS:
        ok = r.array([&] {
          uint32_t m2; uint64_t a2; bool i2;
          if (!r.head(m2, a2, i2)) return false;
          if (m2 != 4) return r.fail("invalid type: %s, expected an array of length 2", CborReader::kind(m2));
          if (i2 || a2 != 2) return r.fail("invalid length, expected an array of length 2");
          uint32_t a = 0, b = 0;
          if (!r.u32(a, "arg_indices") || !r.u32(b, "arg_indices")) return false;
          d.arg0.push_back(a);
          d.arg1.push_back(b);
          return true;
        }, "arg_indices");
        break;
      case F_VALS:
        ok = r.array([&] { double v; if (!r.f64(v)) return false; d.values.push_back(v); return true; }, "values");
        break;
      case F_NIN: ok = r.u32(d.num_inputs, "num_inputs"); break;
      case F_NVAR: ok = r.u32(d.num_variables, "num_variables"); break;
      case F_OUT: ok = r.u32(d.output_index, "output_index"); break;
      case F_OUTS:
        ok = r.array([&] { uint32_t v; if (!r.u32(v, "output_indices")) return false; d.output_indices.push_back(v);
                           return true; }, "output_indices");
        break;
      case F_CSA: {
        uint32_t m2; uint64_t a2; bool i2;
        if (!r.head(m2, a2, i2)) return false;
        if (m2 != 5) return r.fail("invalid type: %s, expected a map (field `custom_second_args`)",
                                   CborReader::kind(m2));
        for (uint64_t q = 0; i2 || q < a2; ++q) {
          if (i2 && r.peek_break()) { ++r.p; break; }
          uint32_t kk, vv;
          if (!r.u32(kk, "custom_second_args") || !r.u32(vv, "custom_second_args")) return false;
          if (d.n_custom_second++ == 0) { d.first_custom_key = kk; d.first_custom_val = vv; }
        }
        break;
      }
    }
    if (!ok) return false;
  }
  for (int f = 0; f <= F_OUT; ++f)
    if (!seen[f]) return r.fail("missing field `%s`", names[f]);
  if (r.p != r.e) return r.fail("trailing data");
  return true;
}

int32_t read_tape_cbor(const uint8_t* buf, uint64_t len, TapeData& d) {
  if (!buf && len) return fail(EHIP_OTHER, "tape CBOR: null buffer");
  CborReader r{buf, buf, buf + len, {}};
  if (!parse_tape_cbor(r, d)) return fail(EHIP_OTHER, "tape CBOR: %s", r.err.c_str());
  std::string msg = check_tape(d);
  if (!msg.empty()) return fail(EHIP_OTHER, "tape CBOR: %s", msg.c_str());
  return EHIP_OK;
}

// The tape's fields after the Custom-op refusal the serializer shares.
int32_t check_writable(const uint32_t* opcodes, uint64_t n, const char* what) {
  for (uint64_t i = 0; i < n; ++i) {
    if (opcodes[i] == OP_CUSTOM_)
      return fail(EHIP_CUSTOM_OPS_NOT_SUPPORTED,
                  "cannot serialize a BytecodeTape containing custom ops; custom ops must be re-registered "
                  "after deserialization");
    if (opcodes[i] >= kNumOps)
      return fail(EHIP_OTHER, "tape %s: opcodes[%llu] (%u) is not an OpCode", what, (unsigned long long)i, opcodes[i]);
  }
  return EHIP_OK;
}

int32_t copy_out(const TapeData& d, int32_t dtype, uint32_t* opcodes, uint32_t* arg0, uint32_t* arg1, void* values,
                 uint64_t cap_entries, uint32_t* output_indices, uint64_t cap_outputs, uint64_t* num_entries,
                 uint32_t* num_inputs, uint32_t* output_index, uint64_t* num_output_indices, const char* what) {
  const uint64_t n = d.opcodes.size(), m = d.output_indices.size();
  *num_entries = n;
  *num_inputs = d.num_inputs;
  *output_index = d.output_index;
  *num_output_indices = m;
  if (!opcodes && !arg0 && !arg1 && !values && !output_indices) return ok();  // sizing call
  if (cap_entries < n || cap_outputs < m || (n && (!opcodes || !arg0 || !arg1 || !values)) ||
      (m && !output_indices))
    return fail(EHIP_OTHER, "tape %s: result buffers too small (%llu entries, %llu outputs)", what,
                (unsigned long long)n, (unsigned long long)m);
  if (n) {
    std::memcpy(opcodes, d.opcodes.data(), n * 4);
    std::memcpy(arg0, d.arg0.data(), n * 4);
    std::memcpy(arg1, d.arg1.data(), n * 4);
  }
  if (dtype == EHIP_F64) {
    if (n) std::memcpy(values, d.values.data(), n * 8);
  } else {
    float* v = (float*)values;
    for (uint64_t i = 0; i < n; ++i) v[i] = (float)d.values[i];
  }
  if (m) std::memcpy(output_indices, d.output_indices.data(), m * 4);
  return ok();
}

}  // namespace

extern "C" {

int32_t ehip_tape_to_json(int32_t dtype, const uint32_t* opcodes, const uint32_t* arg0, const uint32_t* arg1,
                          const void* values, uint64_t num_entries, uint32_t num_inputs, uint32_t output_index,
                          const uint32_t* output_indices, uint64_t num_output_indices, char* buf, uint64_t cap,
                          uint64_t* len) try {
  if (!len) return fail(EHIP_OTHER, "tape JSON: null length pointer");
  if (dtype != EHIP_F32 && dtype != EHIP_F64) return fail(EHIP_OTHER, "tape JSON: dtype must be EHIP_F32 or EHIP_F64");
  if (num_entries && (!opcodes || !arg0 || !arg1 || !values)) return fail(EHIP_OTHER, "tape JSON: null tape array");
  if (num_output_indices && !output_indices) return fail(EHIP_OTHER, "tape JSON: null output_indices");
  if (num_entries > 0xFFFFFFFFull) return fail(EHIP_OTHER, "tape JSON: more than u32::MAX entries");
  int32_t rc = check_writable(opcodes, num_entries, "JSON");
  if (rc) return rc;
  std::string s;
  s.reserve(64 + num_entries * 48);
  s += "{\"opcodes\":[";
  for (uint64_t i = 0; i < num_entries; ++i) {
    if (i) s.push_back(',');
    s.push_back('"');
    s += kOpName[opcodes[i]];
    s.push_back('"');
  }
  s += "],\"arg_indices\":[";
  for (uint64_t i = 0; i < num_entries; ++i) {
    if (i) s.push_back(',');
    s.push_back('[');
    s += std::to_string(arg0[i]);
    s.push_back(',');
    s += std::to_string(arg1[i]);
    s.push_back(']');
  }
  s += "],\"values\":[";
  for (uint64_t i = 0; i < num_entries; ++i) {
    if (i) s.push_back(',');
    if (dtype == EHIP_F64) write_float<double>(s, ((const double*)values)[i]);
    else write_float<float>(s, ((const float*)values)[i]);
  }
  s += "],\"num_inputs\":" + std::to_string(num_inputs);
  s += ",\"num_variables\":" + std::to_string(num_entries);
  s += ",\"output_index\":" + std::to_string(output_index);
  s += ",\"output_indices\":[";
  for (uint64_t j = 0; j < num_output_indices; ++j) {
    if (j) s.push_back(',');
    s += std::to_string(output_indices[j]);
  }
  s += "],\"custom_second_args\":{}}";
  *len = s.size();
  if (!buf) return ok();
  if (cap < s.size() + 1) return fail(EHIP_OTHER, "tape JSON: buffer too small (%llu bytes, need %llu)",
                                      (unsigned long long)cap, (unsigned long long)(s.size() + 1));
  std::memcpy(buf, s.data(), s.size());
  buf[s.size()] = 0;
  return ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_tape_from_json(const char* json, uint64_t len, int32_t dtype, uint32_t* opcodes, uint32_t* arg0,
                            uint32_t* arg1, void* values, uint64_t cap_entries, uint32_t* output_indices,
                            uint64_t cap_outputs, uint64_t* num_entries, uint32_t* num_inputs,
                            uint32_t* output_index, uint64_t* num_output_indices) try {
  if (!num_entries || !num_inputs || !output_index || !num_output_indices)
    return fail(EHIP_OTHER, "tape JSON: null result pointer");
  if (dtype != EHIP_F32 && dtype != EHIP_F64) return fail(EHIP_OTHER, "tape JSON: dtype must be EHIP_F32 or EHIP_F64");
  TapeData d;
  int32_t rc = read_tape(json, len, d);
  if (rc) return rc;
  return copy_out(d, dtype, opcodes, arg0, arg1, values, cap_entries, output_indices, cap_outputs, num_entries,
                  num_inputs, output_index, num_output_indices, "JSON");
} catch (...) {
  return exception_status();
}

int32_t ehip_tape_upload_json(ehip_ctx* ctx, int32_t dtype, const char* json, uint64_t len, ehip_tape** out) try {
  if (!ctx || !out) return fail(EHIP_OTHER, "GPU error: null argument");
  *out = nullptr;
  if (dtype != EHIP_F32 && dtype != EHIP_F64) return fail(EHIP_OTHER, "GPU error: dtype must be EHIP_F32 or EHIP_F64");
  TapeData d;
  int32_t rc = read_tape(json, len, d);
  if (rc) return rc;
  const uint32_t n = (uint32_t)d.opcodes.size();
  std::vector<float> vf;
  const void* vals = d.values.data();
  if (dtype == EHIP_F32) {
    vf.resize(n);
    for (uint32_t i = 0; i < n; ++i) vf[i] = (float)d.values[i];
    vals = vf.data();
  }
  return ehip_tape_upload(ctx, dtype, d.opcodes.data(), d.arg0.data(), d.arg1.data(), vals, n, d.num_inputs, n,
                          d.output_index, d.output_indices.data(), (uint32_t)d.output_indices.size(), out);
} catch (...) {
  return exception_status();
}

int32_t ehip_tape_to_cbor(int32_t dtype, const uint32_t* opcodes, const uint32_t* arg0, const uint32_t* arg1,
                          const void* values, uint64_t num_entries, uint32_t num_inputs, uint32_t output_index,
                          const uint32_t* output_indices, uint64_t num_output_indices, uint8_t* buf, uint64_t cap,
                          uint64_t* len) try {
  if (!len) return fail(EHIP_OTHER, "tape CBOR: null length pointer");
  if (dtype != EHIP_F32 && dtype != EHIP_F64) return fail(EHIP_OTHER, "tape CBOR: dtype must be EHIP_F32 or EHIP_F64");
  if (num_entries && (!opcodes || !arg0 || !arg1 || !values)) return fail(EHIP_OTHER, "tape CBOR: null tape array");
  if (num_output_indices && !output_indices) return fail(EHIP_OTHER, "tape CBOR: null output_indices");
  if (num_entries > 0xFFFFFFFFull) return fail(EHIP_OTHER, "tape CBOR: more than u32::MAX entries");
  int32_t rc = check_writable(opcodes, num_entries, "CBOR");
  if (rc) return rc;
  std::string o;
  o.reserve(64 + num_entries * 24);
  cbor_head(o, 5, 8);
  cbor_text(o, "opcodes");
  cbor_head(o, 4, num_entries);
  for (uint64_t i = 0; i < num_entries; ++i) cbor_text(o, kOpName[opcodes[i]]);
  cbor_text(o, "arg_indices");
  cbor_head(o, 4, num_entries);
  for (uint64_t i = 0; i < num_entries; ++i) {
    cbor_head(o, 4, 2);
    cbor_head(o, 0, arg0[i]);
    cbor_head(o, 0, arg1[i]);
  }
  cbor_text(o, "values");
  cbor_head(o, 4, num_entries);
  for (uint64_t i = 0; i < num_entries; ++i)
    cbor_float(o, dtype == EHIP_F64 ? ((const double*)values)[i] : (double)((const float*)values)[i]);
  cbor_text(o, "num_inputs");
  cbor_head(o, 0, num_inputs);
  cbor_text(o, "num_variables");
  cbor_head(o, 0, num_entries);
  cbor_text(o, "output_index");
  cbor_head(o, 0, output_index);
  cbor_text(o, "output_indices");
  cbor_head(o, 4, num_output_indices);
  for (uint64_t j = 0; j < num_output_indices; ++j) cbor_head(o, 0, output_indices[j]);
  cbor_text(o, "custom_second_args");
  cbor_head(o, 5, 0);
  *len = o.size();
  if (!buf) return ok();
  if (cap < o.size()) return fail(EHIP_OTHER, "tape CBOR: buffer too small (%llu bytes, need %llu)",
                                  (unsigned long long)cap, (unsigned long long)o.size());
  std::memcpy(buf, o.data(), o.size());
  return ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_tape_from_cbor(const uint8_t* cbor, uint64_t len, int32_t dtype, uint32_t* opcodes, uint32_t* arg0,
                            uint32_t* arg1, void* values, uint64_t cap_entries, uint32_t* output_indices,
                            uint64_t cap_outputs, uint64_t* num_entries, uint32_t* num_inputs,
                            uint32_t* output_index, uint64_t* num_output_indices) try {
  if (!num_entries || !num_inputs || !output_index || !num_output_indices)
    return fail(EHIP_OTHER, "tape CBOR: null result pointer");
  if (dtype != EHIP_F32 && dtype != EHIP_F64) return fail(EHIP_OTHER, "tape CBOR: dtype must be EHIP_F32 or EHIP_F64");
  TapeData d;
  int32_t rc = read_tape_cbor(cbor, len, d);
  if (rc) return rc;
  return copy_out(d, dtype, opcodes, arg0, arg1, values, cap_entries, output_indices, cap_outputs, num_entries,
                  num_inputs, output_index, num_output_indices, "CBOR");
} catch (...) {
  return exception_status();
}

}  // extern "C"
