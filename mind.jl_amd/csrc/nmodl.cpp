// NMODL -> HIP C lowering (host code): tokenizer, recursive-descent parser and
// emitter. The interface and the lowering rules are described in nmodl.h.
#include "nmodl.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cctype>
#include <map>
#include <memory>
#include <set>

namespace mind {
namespace nmodl {

// ------------------------------------------------------------------ SHA-256
// FIPS 180-4 SHA-256 of a byte string.
namespace {

struct Sha256 {
  uint32_t h[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a, 0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
  static uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
  void block(const unsigned char* p) {
    static const uint32_t K[64] = {
        0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
        0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
        0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
        0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
        0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
        0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
        0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
        0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};
    uint32_t w[64];
    for (int i = 0; i < 16; ++i)
      w[i] = (uint32_t)p[4 * i] << 24 | (uint32_t)p[4 * i + 1] << 16 | (uint32_t)p[4 * i + 2] << 8 | p[4 * i + 3];
    for (int i = 16; i < 64; ++i) {
      const uint32_t s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
      const uint32_t s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
      w[i] = w[i - 16] + s0 + w[i - 7] + s1;
    }
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], k = h[7];
    for (int i = 0; i < 64; ++i) {
      const uint32_t t1 = k + (rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25)) + ((e & f) ^ (~e & g)) + K[i] + w[i];
      const uint32_t t2 = (rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
      k = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += k;
  }
};

}  // namespace

uint64_t digest(const std::string& text) {
  Sha256 s;
  const size_t n = text.size();
  size_t i = 0;
  for (; i + 64 <= n; i += 64) s.block((const unsigned char*)text.data() + i);
  unsigned char tail[128] = {0};
  const size_t rest = n - i;
  memcpy(tail, text.data() + i, rest);
  tail[rest] = 0x80;
  const size_t len = rest + 9 <= 64 ? 64 : 128;
  const uint64_t bits = (uint64_t)n * 8;
  for (int k = 0; k < 8; ++k) tail[len - 1 - k] = (unsigned char)(bits >> (8 * k));
  s.block(tail);
  if (len == 128) s.block(tail + 64);
  return (uint64_t)s.h[0] << 32 | s.h[1];
}

namespace {

// ------------------------------------------------------------------ text helpers
// Shortest round-trip decimal of x, positional for exponents -4..15, else
// scientific with a signed two-digit exponent (the form Python's repr gives).
std::string float_text(double x) {
  if (isnan(x)) return "nan";
  if (isinf(x)) return x > 0 ? "inf" : "-inf";
  if (x == 0.0) return signbit(x) ? "-0.0" : "0.0";
  char buf[64];
  for (int p = 1; p <= 17; ++p) {
    snprintf(buf, sizeof buf, "%.*e", p - 1, x);
    if (strtod(buf, nullptr) == x) break;
  }
  std::string s = buf, sign;
  if (s[0] == '-') {
    sign = "-";
    s = s.substr(1);
  }
  const size_t epos = s.find('e');
  const int e10 = atoi(s.c_str() + epos + 1);
  std::string digits;
  for (size_t k = 0; k < epos; ++k)
    if (s[k] != '.') digits += s[k];
  while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
  const int nd = (int)digits.size();
  std::string out;
  if (e10 >= -4 && e10 < 16) {
    const int decpt = e10 + 1;
    if (decpt <= 0)
      out = "0." + std::string(-decpt, '0') + digits;
    else if (decpt >= nd)
      out = digits + std::string(decpt - nd, '0') + ".0";
    else
      out = digits.substr(0, decpt) + "." + digits.substr(decpt);
  } else {
    out = digits.substr(0, 1);
    if (nd > 1) out += "." + digits.substr(1);
    char ex[16];
    snprintf(ex, sizeof ex, "e%c%02d", e10 < 0 ? '-' : '+', e10 < 0 ? -e10 : e10);
    out += ex;
  }
  return sign + out;
}

// C++ literal of a double.
std::string lit(double x) {
  const std::string s = float_text(x);
  if (s.find('.') != std::string::npos || s.find('e') != std::string::npos || s.find("inf") != std::string::npos ||
      s.find("nan") != std::string::npos)
    return s;
  return s + ".0";
}

// Quoted form of a name or token in messages.
std::string quoted(const std::string& s) {
  const bool dq = s.find('\'') != std::string::npos && s.find('"') == std::string::npos;
  const char q = dq ? '"' : '\'';
  std::string out(1, q);
  for (unsigned char ch : s) {
    if (ch == '\\') {
      out += "\\\\";
    } else if (ch == (unsigned char)q) {
      out += '\\';
      out += (char)ch;
    } else if (ch < 0x20 || ch == 0x7f) {
      char b[8];
      snprintf(b, sizeof b, "\\x%02x", ch);
      out += b;
    } else {
      out += (char)ch;
    }
  }
  return out + q;
}

std::string hex16(uint64_t v) {
  char b[24];
  snprintf(b, sizeof b, "%016llx", (unsigned long long)v);
  return b;
}

std::string basename_of(const std::string& path) {
  const size_t k = path.rfind('/');
  return k == std::string::npos ? path : path.substr(k + 1);
}

std::string strip(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && isspace((unsigned char)s[a])) ++a;
  while (b > a && isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

bool starts_with(const std::string& s, const char* p) { return s.compare(0, strlen(p), p) == 0; }

[[noreturn]] void raise_at(const std::string& file, int line, const std::string& msg) {
  throw ModError(file + ":" + std::to_string(line) + ": " + msg);
}

// ------------------------------------------------------------------ tokens
struct Tok {
  enum Kind { NAME, NUM, OP, END } kind;
  std::string text;
  int line;
};

bool is_digit(char c) { return c >= '0' && c <= '9'; }
bool name_start(char c) { return (c >= 'A' && c <= 'Z') || (c >= 'a' && c <= 'z') || c == '_'; }
bool name_char(char c) { return name_start(c) || is_digit(c); }

std::vector<Tok> tokenize(const std::string& text, const std::string& file) {
  std::vector<Tok> toks;
  std::vector<std::string> lines;
  size_t start = 0;
  while (true) {
    const size_t nl = text.find('\n', start);
    if (nl == std::string::npos) {
      lines.push_back(text.substr(start));
      break;
    }
    lines.push_back(text.substr(start, nl - start));
    start = nl + 1;
  }
  bool in_comment = false;
  for (size_t li = 0; li < lines.size(); ++li) {
    const int ln = (int)li + 1;
    const std::string& raw = lines[li];
    const std::string stripped = strip(raw);
    if (in_comment) {
      if (starts_with(stripped, "ENDCOMMENT")) in_comment = false;
      continue;
    }
    if (starts_with(stripped, "COMMENT")) {
      in_comment = true;
      continue;
    }
    if (starts_with(stripped, "TITLE")) continue;
    if (starts_with(stripped, "VERBATIM")) raise_at(file, ln, "VERBATIM blocks are not supported");
    std::string line = raw.substr(0, raw.find(':'));  // ':' and '?' start comments
    line = line.substr(0, line.find('?'));
    size_t pos = 0;
    const size_t n = line.size();
    while (pos < n) {
      const char ch = line[pos];
      if (isspace((unsigned char)ch)) {
        ++pos;
        continue;
      }
      const bool dot_num = ch == '.' && pos + 1 < n && is_digit(line[pos + 1]);
      if (is_digit(ch) || dot_num) {  // digits[.digits] or .digits, optional exponent
        size_t j = pos;
        if (is_digit(ch)) {
          while (j < n && is_digit(line[j])) ++j;
          if (j < n && line[j] == '.') {
            ++j;
            while (j < n && is_digit(line[j])) ++j;
          }
        } else {
          j = pos + 1;
          while (j < n && is_digit(line[j])) ++j;
        }
        if (j < n && (line[j] == 'e' || line[j] == 'E')) {
          size_t k = j + 1;
          if (k < n && (line[k] == '+' || line[k] == '-')) ++k;
          if (k < n && is_digit(line[k])) {
            while (k < n && is_digit(line[k])) ++k;
            j = k;
          }
        }
        toks.push_back({Tok::NUM, line.substr(pos, j - pos), ln});
        pos = j;
        continue;
      }
      if (name_start(ch)) {
        size_t j = pos + 1;
        while (j < n && name_char(line[j])) ++j;
        toks.push_back({Tok::NAME, line.substr(pos, j - pos), ln});
        pos = j;
        continue;
      }
      static const char* two[] = {"==", "!=", "<=", ">=", "&&", "||"};
      bool done = false;
      for (const char* op : two) {
        if (line.compare(pos, 2, op) == 0) {
          toks.push_back({Tok::OP, op, ln});
          pos += 2;
          done = true;
          break;
        }
      }
      if (done) continue;
      if (strchr("{}()[]<>,=+-*/^!'~", ch) && ch != '\0') {
        toks.push_back({Tok::OP, std::string(1, ch), ln});
        ++pos;
        continue;
      }
      raise_at(file, ln, "unexpected character " + quoted(std::string(1, ch)));
    }
  }
  if (in_comment) raise_at(file, (int)lines.size(), "unterminated COMMENT");
  toks.push_back({Tok::END, "", (int)lines.size()});
  return toks;
}

// ------------------------------------------------------------------ syntax tree
struct Expr;
using E = std::shared_ptr<const Expr>;
struct Expr {
  enum Kind { NUM, NAME, CALL, UNARY, BINARY, INDEX } kind;
  double num = 0.0;
  std::string id;       // name, called name, operator, or array name
  std::vector<E> args;  // call arguments; {x} of a unary; {a, b} of a binary; {index}
};

E num_expr(double v) { return std::make_shared<Expr>(Expr{Expr::NUM, v, "", {}}); }
E name_expr(const std::string& id) { return std::make_shared<Expr>(Expr{Expr::NAME, 0.0, id, {}}); }
E call_expr(const std::string& id, std::vector<E> args) {
  return std::make_shared<Expr>(Expr{Expr::CALL, 0.0, id, std::move(args)});
}
E unary_expr(const std::string& op, E x) { return std::make_shared<Expr>(Expr{Expr::UNARY, 0.0, op, {x}}); }
E binary_expr(const std::string& op, E a, E b) {
  return std::make_shared<Expr>(Expr{Expr::BINARY, 0.0, op, {a, b}});
}
E index_expr(const std::string& id, E index) { return std::make_shared<Expr>(Expr{Expr::INDEX, 0.0, id, {index}}); }

struct Stmt;
using S = std::shared_ptr<const Stmt>;
struct Stmt {
  enum Kind { ASSIGN, WHILE, IF, CALL, LOCAL, SOLVE } kind;
  std::string target;  // assigned name; SOLVE: the block
  std::string method;  // SOLVE
  E value;             // assigned value; condition of while / if; call of a call statement
  E index;             // element of an array field, or null
  bool derivative = false;
  int line = 0;
  std::vector<S> body, other;
  std::vector<std::string> names;  // LOCAL
};

S assign_stmt(const std::string& target, E value, bool derivative, int line, E index = nullptr) {
  auto s = std::make_shared<Stmt>();
  s->kind = Stmt::ASSIGN;
  s->target = target;
  s->value = value;
  s->derivative = derivative;
  s->line = line;
  s->index = index;
  return s;
}

S call_stmt_of(E call, int line) {
  auto s = std::make_shared<Stmt>();
  s->kind = Stmt::CALL;
  s->value = call;
  s->line = line;
  return s;
}

struct Block {
  std::string kind, name;  // kind: initial, breakpoint, derivative, procedure, function, net_receive
  std::vector<std::string> args;
  std::vector<S> body;
  int line = 0;
};

struct Decl {
  std::string name, storage;  // storage: parameter, state, assigned
  bool has_default = false;
  double value = 0.0;
};

struct Ion {
  std::string ion;
  std::vector<std::string> reads, writes;
  bool has_valence = false;
  double valence = 0.0;
};

// A READ_SOURCE / READ_TARGET / WRITE_SOURCE / WRITE_TARGET binding of a
// MIND-role mechanism: an ROI exposure and the mechanism variable it maps to.
struct Binding {
  std::string direction, exposure, variable;
};

struct Mechanism {
  std::string filename, source, kind, name;  // kind: suffix, point_process, artificial_cell
  std::string role;                          // MIND block ROLE (lower case), empty for micro mechanisms
  std::vector<std::string> exposures;
  std::vector<Binding> bindings;
  std::vector<Ion> ions;
  std::vector<std::string> nonspecific, electrode, ranges, randoms;
  std::vector<std::string> globals;  // GLOBAL names: one value every instance shares
  std::vector<Decl> decls;
  std::vector<Block> blocks;  // in first-definition order; a redefinition replaces the body in place
  std::map<std::string, double> defines;
  std::map<std::string, int> extents;  // array field -> element count

  std::vector<std::string> fields() const {
    std::vector<std::string> order;
    for (const char* storage : {"parameter", "state", "assigned"})
      for (const Decl& d : decls) {
        if (d.storage != storage) continue;
        if (d.name == "v" || d.name == "t" || d.name == "dt" || d.name == "celsius") continue;
        bool have = false;
        for (const std::string& o : order) have |= o == d.name;
        if (!have) order.push_back(d.name);
      }
    return order;
  }
  const Block* block(const std::string& kind, const std::string& name) const {
    for (const Block& b : blocks)
      if (b.kind == kind && b.name == name) return &b;
    return nullptr;
  }
  void set_block(Block b) {
    for (Block& o : blocks)
      if (o.kind == b.kind && o.name == b.name) {
        o = std::move(b);
        return;
      }
    blocks.push_back(std::move(b));
  }
  uint64_t digest() const { return nmodl::digest(source); }
};

// ------------------------------------------------------------------ parser
class Parser {
 public:
  Parser(const std::string& text, const std::string& filename) : toks_(tokenize(text, filename)), file_(filename) {
    m_.filename = filename;
    m_.source = text;
  }

  Mechanism parse() {
    while (peek().kind != Tok::END) {
      const Tok t = take();
      const std::string& k = t.text;
      if (k == "NEURON") {
        neuron();
      } else if (k == "PARAMETER" || k == "ASSIGNED" || k == "STATE") {
        declarations(k == "PARAMETER" ? "parameter" : k == "ASSIGNED" ? "assigned" : "state");
      } else if (k == "UNITS" || k == "INDEPENDENT") {
        skip_block();
      } else if (k == "UNITSOFF" || k == "UNITSON" || k == "THREADSAFE") {
        continue;
      } else if (k == "DEFINE") {
        const std::string name = ident();
        const Tok v = take();
        if (v.kind != Tok::NUM) fail(v, "DEFINE takes a number");
        m_.defines[name] = strtod(v.text.c_str(), nullptr);
      } else if (k == "LOCAL") {
        names_list();
      } else if (k == "MIND") {
        mind_block();
      } else if (k == "INITIAL" || k == "BREAKPOINT" || k == "NET_RECEIVE" || k == "DERIVATIVE" ||
                 k == "PROCEDURE" || k == "FUNCTION") {
        block(k, t);
      } else {
        fail(t, "unexpected top-level token " + quoted(k));
      }
    }
    if (m_.name.empty()) raise_at(file_, 1, "no NEURON block naming the mechanism");
    return m_;
  }

 private:
  std::vector<Tok> toks_;
  size_t i_ = 0;
  std::string file_;
  Mechanism m_;

  const Tok& peek(size_t k = 0) const { return toks_[std::min(i_ + k, toks_.size() - 1)]; }
  Tok take() {
    const Tok t = peek();
    ++i_;
    return t;
  }
  bool at(const char* text) const {
    const Tok& t = peek();
    return (t.kind == Tok::OP || t.kind == Tok::NAME) && t.text == text;
  }
  bool accept(const char* text) {
    if (at(text)) {
      ++i_;
      return true;
    }
    return false;
  }
  Tok expect(const char* text) {
    const Tok t = take();
    if (t.text != text) fail(t, std::string("expected ") + quoted(text) + ", found " + quoted(t.text));
    return t;
  }
  std::string ident() {
    const Tok t = take();
    if (t.kind != Tok::NAME) fail(t, "expected a name, found " + quoted(t.text));
    return t.text;
  }
  [[noreturn]] void fail(const Tok& t, const std::string& msg) const { raise_at(file_, t.line, msg); }

  void skip_parens() {
    int depth = 0;
    while (true) {
      const Tok t = take();
      if (t.kind == Tok::END) fail(t, "unterminated parentheses");
      if (t.text == "(") {
        ++depth;
      } else if (t.text == ")") {
        if (--depth == 0) return;
      }
    }
  }
  void skip_block() {
    expect("{");
    int depth = 1;
    while (depth) {
      const Tok t = take();
      if (t.kind == Tok::END) fail(t, "unterminated block");
      if (t.text == "{") ++depth;
      if (t.text == "}") --depth;
    }
  }
  std::vector<std::string> names_list() {
    std::vector<std::string> names{ident()};
    while (accept(",")) names.push_back(ident());
    return names;
  }
  static void append(std::vector<std::string>& to, const std::vector<std::string>& from) {
    to.insert(to.end(), from.begin(), from.end());
  }

  void neuron() {
    expect("{");
    while (!accept("}")) {
      const Tok t = take();
      const std::string& k = t.text;
      if (k == "SUFFIX" || k == "POINT_PROCESS" || k == "ARTIFICIAL_CELL") {
        m_.kind = k == "SUFFIX" ? "suffix" : k == "POINT_PROCESS" ? "point_process" : "artificial_cell";
        m_.name = ident();
      } else if (k == "USEION") {
        Ion ion;
        ion.ion = ident();
        while (peek().text == "READ" || peek().text == "WRITE" || peek().text == "VALENCE") {
          const std::string w = take().text;
          if (w == "VALENCE") {
            const bool neg = accept("-");
            const Tok v = take();
            if (v.kind != Tok::NUM) fail(v, "VALENCE takes a number");
            ion.has_valence = true;
            ion.valence = neg ? -strtod(v.text.c_str(), nullptr) : strtod(v.text.c_str(), nullptr);
            continue;
          }
          append(w == "READ" ? ion.reads : ion.writes, names_list());
        }
        m_.ions.push_back(ion);
      } else if (k == "NONSPECIFIC_CURRENT") {
        append(m_.nonspecific, names_list());
      } else if (k == "ELECTRODE_CURRENT") {
        append(m_.electrode, names_list());
      } else if (k == "RANGE" || k == "GLOBAL") {
        const std::vector<std::string> names = names_list();
        append(m_.ranges, names);
        if (k == "GLOBAL") append(m_.globals, names);
      } else if (k == "RANDOM") {
        append(m_.randoms, names_list());
      } else if (k == "THREADSAFE") {
        continue;
      } else if (k == "POINTER" || k == "BBCOREPOINTER") {
        fail(t, k + " is not supported by the device lowering");
      } else {
        fail(t, "unsupported NEURON statement " + quoted(k));
      }
    }
  }

  // MIND { ROLE r  EXPOSURE a, b  READ_SOURCE x AS y, ... } (frontend.jl:514-548)
  void mind_block() {
    expect("{");
    auto keyword = [](const std::string& w) {
      return w == "ROLE" || w == "EXPOSURE" || w == "READ_SOURCE" || w == "READ_TARGET" || w == "WRITE_SOURCE" ||
             w == "WRITE_TARGET";
    };
    while (!accept("}")) {
      const Tok t = take();
      if (t.text == "ROLE") {
        std::string role = ident();
        for (char& ch : role)
          if (ch >= 'A' && ch <= 'Z') ch = (char)(ch - 'A' + 'a');
        if (role != "region" && role != "macro2macro" && role != "macro2micro" && role != "micro2macro")
          fail(t, "unknown MIND role " + role);
        m_.role = role;
      } else if (t.text == "EXPOSURE") {
        append(m_.exposures, names_list());
      } else if (t.text == "READ_SOURCE" || t.text == "READ_TARGET" || t.text == "WRITE_SOURCE" ||
                 t.text == "WRITE_TARGET") {
        while (peek().kind == Tok::NAME && !keyword(peek().text)) {
          Binding b;
          b.direction = t.text;
          b.exposure = ident();
          b.variable = accept("AS") ? ident() : b.exposure;
          m_.bindings.push_back(b);
          accept(",");
        }
      } else {
        fail(t, "unknown MIND declaration " + t.text);
      }
    }
  }

  void declarations(const std::string& storage) {
    expect("{");
    while (!accept("}")) {
      Decl d;
      d.name = ident();
      d.storage = storage;
      if (accept("[")) {  // array field: name[N], N a number or a DEFINE
        const Tok t = take();
        bool known = true;
        double n = 0.0;
        if (t.kind == Tok::NUM) {
          n = strtod(t.text.c_str(), nullptr);
        } else {
          auto it = m_.defines.find(t.text);
          known = it != m_.defines.end();
          if (known) n = it->second;
        }
        if (!known || n != floor(n) || n < 1) fail(t, "array size must be a positive integer or DEFINE");
        expect("]");
        m_.extents[d.name] = (int)n;
      }
      if (accept("=")) {
        const bool neg = accept("-");
        const Tok t = take();
        if (t.kind != Tok::NUM) fail(t, "declaration defaults must be numbers");
        d.has_default = true;
        d.value = neg ? -strtod(t.text.c_str(), nullptr) : strtod(t.text.c_str(), nullptr);
      }
      if (at("(")) skip_parens();
      if (accept("<")) {
        while (!accept(">"))
          if (take().kind == Tok::END) fail(peek(), "unterminated bounds");
      }
      m_.decls.push_back(d);
    }
  }

  void block(const std::string& k, const Tok& tok) {
    Block b;
    b.kind = k == "NET_RECEIVE" ? "net_receive" : k == "INITIAL"      ? "initial"
                                              : k == "BREAKPOINT"     ? "breakpoint"
                                              : k == "DERIVATIVE"     ? "derivative"
                                              : k == "PROCEDURE"      ? "procedure"
                                                                      : "function";
    b.name = b.kind;
    if (b.kind == "derivative" || b.kind == "procedure" || b.kind == "function") b.name = ident();
    if (accept("(")) {
      while (!accept(")")) {
        b.args.push_back(ident());
        if (at("(")) skip_parens();
        accept(",");
      }
    }
    if (at("(")) skip_parens();  // FUNCTION result units
    b.body = body();
    b.line = tok.line;
    m_.set_block(std::move(b));
  }

  std::vector<S> body() {
    expect("{");
    std::vector<S> out;
    while (!accept("}")) {
      S s = statement();
      if (s) out.push_back(s);
    }
    return out;
  }

  S statement() {
    const Tok t = peek();
    if (t.kind == Tok::END) fail(t, "unterminated block");
    if (accept("LOCAL")) {
      auto s = std::make_shared<Stmt>();
      s->kind = Stmt::LOCAL;
      s->names = names_list();
      return s;
    }
    if (accept("SOLVE")) {
      auto s = std::make_shared<Stmt>();
      s->kind = Stmt::SOLVE;
      s->target = ident();
      s->method = "cnexp";
      if (accept("METHOD")) s->method = ident();
      if (s->method != "cnexp" && s->method != "euler" && s->method != "after_cvode")
        fail(t, "SOLVE METHOD " + s->method + " is not supported");
      s->line = t.line;
      return s;
    }
    if (accept("TABLE")) {  // tables only memoise; the lowering evaluates the expressions directly
      while (!at("WITH"))
        if (take().kind == Tok::END) fail(t, "unterminated TABLE");
      take();
      take();
      return nullptr;
    }
    if (accept("UNITSOFF") || accept("UNITSON")) return nullptr;
    if (accept("while")) {
      expect("(");
      auto s = std::make_shared<Stmt>();
      s->kind = Stmt::WHILE;
      s->value = expr();
      expect(")");
      s->body = body();
      s->line = t.line;
      return s;
    }
    if (accept("if")) {
      expect("(");
      auto s = std::make_shared<Stmt>();
      s->kind = Stmt::IF;
      s->value = expr();
      expect(")");
      s->body = body();
      if (accept("else")) {
        if (at("if"))
          s->other = {statement()};
        else
          s->other = body();
      }
      s->line = t.line;
      return s;
    }
    const std::string name = ident();
    if (accept("[")) {
      E index = expr();
      expect("]");
      expect("=");
      return assign_stmt(name, expr(), false, t.line, index);
    }
    if (accept("'")) {
      expect("=");
      return assign_stmt(name, expr(), true, t.line);
    }
    if (accept("=")) return assign_stmt(name, expr(), false, t.line);
    if (at("(")) return call_stmt_of(call_args(name), t.line);
    fail(t, "expected an assignment or a call after " + quoted(name));
  }

  E call_args(const std::string& name) {
    expect("(");
    std::vector<E> args;
    while (!accept(")")) {
      args.push_back(expr());
      accept(",");
    }
    return call_expr(name, args);
  }

  // || < && < comparisons < + - < * / < unary < ^
  E expr() {
    E a = conj();
    while (accept("||")) a = binary_expr("||", a, conj());
    return a;
  }
  E conj() {
    E a = comparison();
    while (accept("&&")) a = binary_expr("&&", a, comparison());
    return a;
  }
  E comparison() {
    E a = additive();
    while (true) {
      const std::string& p = peek().text;
      if (p != "<" && p != ">" && p != "<=" && p != ">=" && p != "==" && p != "!=") break;
      const std::string op = take().text;
      a = binary_expr(op, a, additive());
    }
    return a;
  }
  E additive() {
    E a = term();
    while ((peek().text == "+" || peek().text == "-") && peek().kind == Tok::OP) {
      const std::string op = take().text;
      a = binary_expr(op, a, term());
    }
    return a;
  }
  E term() {
    E a = unary();
    while ((peek().text == "*" || peek().text == "/") && peek().kind == Tok::OP) {
      const std::string op = take().text;
      a = binary_expr(op, a, unary());
    }
    return a;
  }
  E unary() {
    if (accept("-")) return unary_expr("-", unary());
    if (accept("!")) return unary_expr("!", unary());
    return power();
  }
  E power() {
    E a = primary();
    if (accept("^")) return binary_expr("^", a, unary());
    return a;
  }
  E primary() {
    const Tok t = take();
    if (t.kind == Tok::NUM) return num_expr(strtod(t.text.c_str(), nullptr));
    if (t.text == "(") {
      E e = expr();
      expect(")");
      return e;
    }
    if (t.kind == Tok::NAME) {
      if (at("(")) return call_args(t.text);
      if (accept("[")) {
        E index = expr();
        expect("]");
        return index_expr(t.text, index);
      }
      return name_expr(t.text);
    }
    fail(t, "unexpected " + quoted(t.text) + " in expression");
  }
};

Mechanism parse(const std::string& text, const std::string& filename) { return Parser(text, filename).parse(); }

// ------------------------------------------------------------------ emitter
using Scope = std::map<std::string, std::string>;
using Chain = std::vector<std::string>;

bool in(const Chain& c, const std::string& s) {
  for (const std::string& x : c)
    if (x == s) return true;
  return false;
}
Chain plus(Chain c, const std::string& s) {
  c.push_back(s);
  return c;
}
std::string join(const std::vector<std::string>& parts, const std::string& sep) {
  std::string out;
  for (size_t k = 0; k < parts.size(); ++k) out += (k ? sep : "") + parts[k];
  return out;
}
bool is_binary_not_pow(const E& x) { return x->kind == Expr::BINARY && x->id != "^"; }

const std::map<std::string, std::string>& elementary() {
  static const std::map<std::string, std::string> f = {
      {"exp", "mind_exp"}, {"log", "mind_log"}, {"fabs", "mind_fabs"}, {"sqrt", "sqrt"}};
  return f;
}

class Gen {
 public:
  Gen(const Mechanism& m, const std::string& prefix) : m_(m), p_(prefix) {
    fields_ = m.fields();
    int off = 0;
    for (const std::string& n : fields_) {
      findex_[n] = off;
      auto it = m.extents.find(n);
      off += it == m.extents.end() ? 1 : it->second;
    }
    ncols_ = off;
    for (const Block& b : m.blocks) {
      if (b.kind == "function") funcs_.push_back(&b);
      if (b.kind == "procedure") procs_.push_back(&b);
      if (b.kind == "derivative") derivs_.push_back(&b);
    }
  }

  int ncols() const { return ncols_; }
  const std::vector<std::string>& fields() const { return fields_; }
  int column(const std::string& n) const { return findex_.at(n); }
  // Fields a block assigns, through the routines it calls.
  std::vector<std::string> block_writes(const Block& b) const {
    std::vector<std::string> acc;
    std::set<std::string> local(b.args.begin(), b.args.end());
    if (b.kind == "function") local.insert(b.name);
    collect_writes(b.body, acc, {b.name}, local);
    return acc;
  }

  std::string emit() {
    std::vector<std::string> out;
    out.push_back("// ---- " + m_.name + " (" + basename_of(m_.filename) + ", sha256 " + hex16(m_.digest()) +
                  ") ----");
    std::vector<std::string> en;
    for (const std::string& n : fields_) en.push_back(p_ + "_" + n + " = " + std::to_string(findex_[n]));
    en.push_back(p_ + "_NFIELDS = " + std::to_string(ncols_));
    out.push_back("enum { " + join(en, ", ") + " };");
    if (!m_.randoms.empty()) out.push_back("constexpr uint64_t " + p_ + "_digest = 0x" + hex16(m_.digest()) + "ull;");
    std::map<std::string, double> defaults;
    for (const Decl& d : m_.decls)
      if (d.storage == "parameter" && d.has_default) defaults[d.name] = d.value;
    out.push_back("MIND_HD void " + p_ + "_defaults(double* f) {");
    for (const std::string& n : fields_)
      if (defaults.count(n)) out.push_back("  f[" + p_ + "_" + n + "] = " + lit(defaults[n]) + ";");
    out.push_back("}");
    std::vector<const Block*> routines = funcs_;
    routines.insert(routines.end(), procs_.begin(), procs_.end());
    for (const Block* b : routines) out.push_back(signature(*b) + ";");
    for (const Block* b : routines) routine(*b, out);
    if (m_.role.empty())
      current(out);
    else
      evaluate(out);
    advance(out);
    initial(out);
    receive(out);
    return join(out, "\n");
  }

 private:
  const Mechanism& m_;
  std::string p_;
  std::vector<std::string> fields_;
  std::map<std::string, int> findex_;
  int ncols_ = 0;
  int tmp_ = 0;
  std::vector<const Block*> funcs_, procs_, derivs_;

  [[noreturn]] void err(const std::string& msg, int line = 1) const { raise_at(m_.filename, line, msg); }
  const Block* find(const std::vector<const Block*>& v, const std::string& name) const {
    for (const Block* b : v)
      if (b->name == name) return b;
    return nullptr;
  }
  const Block* func(const std::string& n) const { return find(funcs_, n); }
  const Block* proc(const std::string& n) const { return find(procs_, n); }
  const Block* deriv(const std::string& n) const { return find(derivs_, n); }
  int extent(const std::string& n) const {
    auto it = m_.extents.find(n);
    return it == m_.extents.end() ? 0 : it->second;
  }
  bool is_random(const std::string& n) const {
    for (const std::string& r : m_.randoms)
      if (r == n) return true;
    return false;
  }

  // expression with a scope: name -> C expression
  std::string e(const E& x, const Scope& scope) const {
    switch (x->kind) {
      case Expr::NUM: return lit(x->num);
      case Expr::NAME: {
        auto it = scope.find(x->id);
        if (it != scope.end()) return it->second;
        err("unknown name " + x->id);
      }
      case Expr::UNARY: return x->id == "!" ? "(!" + e(x->args[0], scope) + ")" : "(-" + e(x->args[0], scope) + ")";
      case Expr::BINARY: {
        const std::string a = e(x->args[0], scope), b = e(x->args[1], scope);
        if (x->id == "^") return "mind_pow(" + a + ", " + b + ")";
        return "(" + a + " " + x->id + " " + b + ")";
      }
      case Expr::CALL: return call(x, scope);
      case Expr::INDEX: return element(x->id, x->args[0], scope);
    }
    err("unsupported expression");
  }

  // Column of element `index` (0-based, as in MOD) of an array field. A constant
  // index is checked here; a computed one is clamped to the array (gen_index).
  std::string element(const std::string& name, const E& index, const Scope& scope) const {
    const int n = extent(name);
    if (n == 0 || scope.count(name)) err(name + " is not an array field");
    if (index->kind == Expr::NUM) {
      const double k = index->num;
      if (k != floor(k) || !(0 <= k && k < n))
        err("index " + float_text(k) + " is outside " + name + "[" + std::to_string(n) + "]");
      return "c(" + p_ + "_" + name + " + " + std::to_string((long long)k) + ", r)";
    }
    return "c(" + p_ + "_" + name + " + gen_index(" + e(index, scope) + ", " + std::to_string(n) + "), r)";
  }

  std::string call(const E& c, const Scope& scope) const {
    std::vector<std::string> args;
    for (const E& a : c->args) args.push_back(e(a, scope));
    auto f = elementary().find(c->id);
    if (f != elementary().end()) {
      if (args.size() != 1) err(c->id + " takes one argument");
      return f->second + "(" + args[0] + ")";
    }
    if (c->id == "random_negexp") {
      need_random(c);
      return "(-mind_log(stream_uniform(*x.s, x.seed)) * 1.0)";
    }
    if (c->id == "random_uniform") {
      need_random(c);
      return "stream_uniform(*x.s, x.seed)";
    }
    if ((c->id == "nernst" || c->id == "nrn_nernst") && !func(c->id)) {  // codegen.jl:112-115
      if (args.size() != 3) err(c->id + " takes (internal, external, valence)");
      return "gen_nernst(" + args[0] + ", " + args[1] + ", " + args[2] + ", x.celsius)";
    }
    if ((c->id == "ghk" || c->id == "nrn_ghk") && !func(c->id)) {
      if (args.size() != 4) err(c->id + " takes (v, internal, external, valence)");
      return "gen_ghk(" + args[0] + ", " + args[1] + ", " + args[2] + ", " + args[3] + ", x.celsius)";
    }
    if (c->id == "at_time") {  // 1 in the step whose interval (t - dt, t] holds the event (codegen.jl:102-110)
      if (args.size() != 1) err("at_time takes one argument");
      return "(((" + args[0] + " - 1.0e-11) <= x.t && (" + args[0] + " - 1.0e-11) > x.t - x.dt) ? 1.0 : 0.0)";
    }
    if (const Block* b = func(c->id)) {
      if (args.size() != b->args.size()) err(c->id + " takes " + std::to_string(b->args.size()) + " arguments", b->line);
      std::string s = p_ + "_" + c->id + "(c, r, x";
      for (const std::string& a : args) s += ", " + a;
      return s + ")";
    }
    err("unknown function " + c->id);
  }

  void need_random(const E& c) const {
    if (c->args.empty() || c->args[0]->kind != Expr::NAME || !is_random(c->args[0]->id))
      err(c->id + " needs a RANDOM variable");
  }

  Scope base_scope() const {
    Scope s;
    for (const auto& d : m_.defines) s[d.first] = lit(d.second);
    for (const std::string& n : fields_)
      if (!m_.extents.count(n)) s[n] = "c(" + p_ + "_" + n + ", r)";
    s["v"] = "x.v";
    s["t"] = "x.t";
    s["dt"] = "x.dt";
    s["celsius"] = "x.celsius";
    for (const std::string& rv : m_.randoms) s[rv] = "0.0";
    return s;
  }

  // Statements of `body` into `out`. `chain` is the chain of routines being
  // expanded in place (the current phase works on local copies of the fields it
  // writes, so its procedure and function calls are expanded instead of called).
  Scope stmts(const std::vector<S>& body, Scope scope, std::vector<std::string>& out, const std::string& ind,
              const char* deriv_method = nullptr, const Chain& chain = {}) {
    const bool inl = !chain.empty();
    for (const S& st : body) {
      switch (st->kind) {
        case Stmt::LOCAL:
          for (const std::string& n : st->names) {
            const std::string cn = "l_" + n;
            out.push_back(ind + "double " + cn + " = 0.0;");
            scope[n] = cn;
          }
          break;
        case Stmt::ASSIGN:
          if (st->index) {
            if (inl) err("array element writes in BREAKPOINT's current evaluation are not supported", st->line);
            const std::string el = element(st->target, st->index, scope);
            out.push_back(ind + el + " = " + e(st->value, scope) + ";");
            break;
          } else {
            auto it = scope.find(st->target);
            if (it == scope.end() || it->second == "x.v" || it->second == "x.t" || it->second == "x.dt" ||
                it->second == "x.celsius")
              err("cannot assign " + st->target, st->line);
            const std::string tgt = it->second;
            E value = st->value;
            if (inl) value = hoist(value, scope, out, ind, chain);
            if (st->derivative)
              out.push_back(ind + derivative(st->target, value, st->line, scope, tgt, deriv_method));
            else
              out.push_back(ind + tgt + " = " + e(value, scope) + ";");
          }
          break;
        case Stmt::WHILE: {
          if (inl && !user_calls(st->value).empty())
            err("a FUNCTION call in a while condition of the current evaluation", st->line);
          std::string cond = e(st->value, scope);
          if (is_binary_not_pow(st->value)) cond = cond.substr(1, cond.size() - 2);
          out.push_back(ind + "while (" + cond + ") {");
          stmts(st->body, scope, out, ind + "  ", deriv_method, chain);
          out.push_back(ind + "}");
          break;
        }
        case Stmt::IF: {
          std::string cond = e(inl ? hoist(st->value, scope, out, ind, chain) : st->value, scope);
          if (is_binary_not_pow(st->value)) cond = cond.substr(1, cond.size() - 2);  // the if's own parentheses
          out.push_back(ind + "if (" + cond + ") {");
          stmts(st->body, scope, out, ind + "  ", deriv_method, chain);
          if (!st->other.empty()) {
            out.push_back(ind + "} else {");
            stmts(st->other, scope, out, ind + "  ", deriv_method, chain);
          }
          out.push_back(ind + "}");
          break;
        }
        case Stmt::CALL: {
          const E& c = st->value;
          if (inl && proc(c->id)) {
            std::vector<E> args;
            for (const E& a : c->args) args.push_back(hoist(a, scope, out, ind, chain));
            inline_proc(call_expr(c->id, args), st->line, scope, out, ind, chain);
          } else if (inl && func(c->id)) {
            hoist(c, scope, out, ind, chain);  // value unused
          } else if (inl) {
            std::vector<E> args;
            for (const E& a : c->args) args.push_back(hoist(a, scope, out, ind, chain));
            out.push_back(ind + call_stmt(call_expr(c->id, args), st->line, scope) + ";");
          } else {
            out.push_back(ind + call_stmt(c, st->line, scope) + ";");
          }
          break;
        }
        case Stmt::SOLVE: err("SOLVE outside BREAKPOINT", st->line);
      }
    }
    return scope;
  }

  // FUNCTION calls in expression x, arguments before the call.
  std::vector<E> user_calls(const E& x) const {
    std::vector<E> out;
    switch (x->kind) {
      case Expr::CALL:
        for (const E& a : x->args) {
          std::vector<E> inner = user_calls(a);
          out.insert(out.end(), inner.begin(), inner.end());
        }
        if (func(x->id)) out.push_back(x);
        break;
      case Expr::UNARY: return user_calls(x->args[0]);
      case Expr::BINARY: {
        out = user_calls(x->args[0]);
        std::vector<E> b = user_calls(x->args[1]);
        out.insert(out.end(), b.begin(), b.end());
        break;
      }
      case Expr::INDEX: return user_calls(x->args[0]);
      default: break;
    }
    return out;
  }

  // In the current evaluation a FUNCTION may assign fields (held in local copies
  // there), so its calls are expanded before the statement that uses them, in
  // evaluation order, each into a temporary that replaces the call. Calls on the
  // right of && or || would lose their short-circuit and are refused.
  E hoist(const E& x, Scope& scope, std::vector<std::string>& out, const std::string& ind, const Chain& chain) {
    switch (x->kind) {
      case Expr::CALL: {
        std::vector<E> args;
        for (const E& a : x->args) args.push_back(hoist(a, scope, out, ind, chain));
        if (!func(x->id)) return call_expr(x->id, args);
        return name_expr(inline_func(call_expr(x->id, args), scope, out, ind, chain));
      }
      case Expr::UNARY: return unary_expr(x->id, hoist(x->args[0], scope, out, ind, chain));
      case Expr::BINARY: {
        if ((x->id == "&&" || x->id == "||") && !user_calls(x->args[1]).empty())
          err("a FUNCTION call on the right of " + x->id + " in the current evaluation");
        E a = hoist(x->args[0], scope, out, ind, chain);
        E b = hoist(x->args[1], scope, out, ind, chain);
        return binary_expr(x->id, a, b);
      }
      case Expr::INDEX: return index_expr(x->id, hoist(x->args[0], scope, out, ind, chain));
      default: return x;
    }
  }

  std::string inline_func(const E& c, Scope& scope, std::vector<std::string>& out, const std::string& ind,
                          const Chain& chain) {
    const Block& f = *func(c->id);
    if (in(chain, f.name)) err("FUNCTION " + f.name + " calls itself", f.line);
    if (c->args.size() != f.args.size()) err(f.name + " takes " + std::to_string(f.args.size()) + " arguments", f.line);
    ++tmp_;
    const std::string tmp = "f_" + f.name + "_" + std::to_string(tmp_);
    out.push_back(ind + "double " + tmp + " = 0.0;");
    out.push_back(ind + "{  // FUNCTION " + f.name);
    Scope inner = scope;
    for (size_t k = 0; k < f.args.size(); ++k) {
      const std::string pn = "p_" + f.name + "_" + f.args[k] + "_" + std::to_string(tmp_);
      out.push_back(ind + "  const double " + pn + " = " + e(c->args[k], scope) + ";");
      inner[f.args[k]] = pn;
    }
    inner[f.name] = tmp;
    stmts(f.body, inner, out, ind + "  ", nullptr, plus(chain, f.name));
    out.push_back(ind + "}");
    scope[tmp] = tmp;
    return tmp;
  }

  void inline_proc(const E& c, int line, Scope& scope, std::vector<std::string>& out, const std::string& ind,
                   const Chain& chain) {
    const Block& p = *proc(c->id);
    if (in(chain, p.name)) err("PROCEDURE " + p.name + " calls itself", line);
    if (c->args.size() != p.args.size()) err(p.name + " takes " + std::to_string(p.args.size()) + " arguments", line);
    out.push_back(ind + "{  // PROCEDURE " + p.name);
    Scope inner = scope;
    for (size_t k = 0; k < p.args.size(); ++k) {
      const std::string pn = "p_" + p.name + "_" + p.args[k];
      out.push_back(ind + "  const double " + pn + " = " + e(c->args[k], scope) + ";");
      inner[p.args[k]] = pn;
    }
    stmts(p.body, inner, out, ind + "  ", nullptr, plus(chain, p.name));
    out.push_back(ind + "}");
  }

  std::string call_stmt(const E& c, int line, const Scope& scope) const {
    std::vector<std::string> args;
    for (const E& a : c->args) args.push_back(e(a, scope));
    if (c->id == "net_send") {
      if (args.size() != 2) err("net_send takes a delay and a flag", line);
      return "x.out.send = 1, x.out.send_delay = " + args[0] + ", x.out.send_flag = " + args[1];
    }
    if (c->id == "net_event") {
      if (args.size() != 1) err("net_event takes a time", line);
      return "x.out.spiked = 1, x.out.spike_time = " + args[0];
    }
    if (c->id == "random_setseq") {
      need_random(c);
      if (args.size() != 2) err("random_setseq takes a RANDOM variable and a sequence", line);
      return "stream_setseq(*x.s, x.seed, " + args[1] + ")";
    }
    if (c->id == "at_time") return "(void)0";  // as a statement its value is unused
    if (const Block* p = proc(c->id)) {
      if (args.size() != p->args.size()) err(c->id + " takes " + std::to_string(p->args.size()) + " arguments", line);
      std::string s = p_ + "_" + c->id + "(c, r, x";
      for (const std::string& a : args) s += ", " + a;
      return s + ")";
    }
    if (func(c->id)) return call(c, scope);
    err("unknown procedure " + c->id, line);
  }

  std::string derivative(const std::string& name, const E& v, int line, const Scope& scope, const std::string& tgt,
                         const char* method) const {
    const std::string dt = scope.at("dt");
    if (method && std::string(method) == "euler")
      return tgt + " = (" + tgt + " + (" + dt + " * (" + e(v, scope) + ")));";
    // cnexp relaxation and pure decay (codegen.jl:599-626)
    if (v->kind == Expr::BINARY && v->id == "/") {
      const E& num = v->args[0];
      const E& tau = v->args[1];
      if (tau->kind == Expr::NUM || (tau->kind == Expr::NAME && tau->id != name)) {
        bool found = false;
        std::string eq;
        if (num->kind == Expr::BINARY && num->id == "-" && num->args[1]->kind == Expr::NAME &&
            num->args[1]->id == name) {
          const E& a = num->args[0];
          if (a->kind == Expr::NUM || (a->kind == Expr::NAME && a->id != name)) {
            eq = e(a, scope);
            found = true;
          }
        } else if (num->kind == Expr::UNARY && num->id == "-" && num->args[0]->kind == Expr::NAME &&
                   num->args[0]->id == name) {
          eq = "0.0";
          found = true;
        }
        if (found) {
          const std::string scale = e(tau, scope);
          return tgt + " = (" + tgt + " + ((1.0 - mind_exp((" + dt + " * (-1.0 / " + scale + ")))) * (" + eq +
                 " - " + tgt + ")));";
        }
      }
    }
    err("cnexp form of " + name + "' is outside the relaxation / decay forms this lowering supports", line);
  }

  std::string signature(const Block& b) const {
    std::string args;
    for (const std::string& a : b.args) args += ", double a_" + a;
    return std::string("MIND_HD ") + (b.kind == "function" ? "double" : "void") + " " + p_ + "_" + b.name +
           "(const Cols& c, int64_t r, GenCtx& x" + args + ")";
  }

  void routine(const Block& b, std::vector<std::string>& out) {
    out.push_back(signature(b) + " {");
    Scope scope = base_scope();
    for (const std::string& a : b.args) {
      out.push_back("  double p_" + a + " = a_" + a + ";");
      out.push_back("  (void)p_" + a + ";");
      scope[a] = "p_" + a;
    }
    if (b.kind == "function") {
      out.push_back("  double f_" + b.name + " = 0.0;");
      scope[b.name] = "f_" + b.name;
    }
    stmts(b.body, scope, out, "  ");
    if (b.kind == "function") out.push_back("  return f_" + b.name + ";");
    out.push_back("}");
  }

  // Fields the current phase assigns, in first-assignment order, through the
  // routines it calls (expanded in place).
  void collect_writes(const std::vector<S>& body, std::vector<std::string>& acc, const Chain& chain,
                      std::set<std::string> local) const {
    for (const S& st : body) {
      if (st->kind == Stmt::LOCAL) {
        local.insert(st->names.begin(), st->names.end());
      } else if (st->kind == Stmt::ASSIGN && findex_.count(st->target) && !local.count(st->target) &&
                 !in(acc, st->target)) {
        acc.push_back(st->target);
      } else if (st->kind == Stmt::IF) {
        collect_calls(st->value, acc, chain);
        collect_writes(st->body, acc, chain, local);
        collect_writes(st->other, acc, chain, local);
      } else if (st->kind == Stmt::WHILE) {
        collect_writes(st->body, acc, chain, local);
      } else if (st->kind == Stmt::CALL && proc(st->value->id)) {
        const Block& p = *proc(st->value->id);
        if (in(chain, p.name)) err("PROCEDURE " + p.name + " calls itself", st->line);
        for (const E& a : st->value->args) collect_calls(a, acc, chain);
        collect_writes(p.body, acc, plus(chain, p.name), std::set<std::string>(p.args.begin(), p.args.end()));
      } else if (st->kind == Stmt::CALL) {
        collect_calls(st->value, acc, chain);
      }
      if (st->kind == Stmt::ASSIGN) collect_calls(st->value, acc, chain);
    }
  }

  void collect_calls(const E& x, std::vector<std::string>& acc, const Chain& chain) const {
    for (const E& c : user_calls(x)) {
      const Block& f = *func(c->id);
      if (in(chain, f.name)) err("FUNCTION " + f.name + " calls itself", f.line);
      std::set<std::string> local(f.args.begin(), f.args.end());
      local.insert(f.name);
      collect_writes(f.body, acc, plus(chain, f.name), local);
    }
  }

  // nrn_cur in two parts, so that a caller can issue the column reads of several
  // rows before any row's stores: <P>_current_load reads every column the
  // evaluation reads into <P>_cur_in; <P>_current_eval evaluates BREAKPOINT at
  // v + 0.001 then v from those values and stores the assigned variables.
  void current(std::vector<std::string>& out) {
    const Block* bp = m_.block("breakpoint", "breakpoint");
    std::vector<S> body;
    if (bp)
      for (const S& s : bp->body)
        if (s->kind != Stmt::SOLVE) body.push_back(s);
    std::vector<std::string> written;
    collect_writes(body, written, {}, {});
    std::vector<std::string> currents = m_.nonspecific;
    std::vector<std::string> ion_currents;  // written ion currents in USEION order (the ish outputs)
    for (const Ion& ion : m_.ions)
      for (const std::string& w : ion.writes)
        if (w == "i" + ion.ion) {
          currents.push_back(w);
          ion_currents.push_back(w);
        }
    if (ion_currents.size() > (size_t)kIshSlots)
      err("a mechanism writes more than " + std::to_string(kIshSlots) + " ion currents");
    std::vector<std::string> all = currents;
    all.insert(all.end(), m_.electrode.begin(), m_.electrode.end());
    for (const std::string& n : all)
      if (!findex_.count(n)) err("current " + n + " is not declared");
    std::vector<std::string> lines;
    lines.push_back("  for (int probe = 0; probe < 2; ++probe) {");
    lines.push_back("    GenCtx x = gen_ctx(probe == 0 ? v + 0.001 : v, t, dt, celsius);");
    lines.push_back("    (void)x;");
    Scope scope = base_scope();
    for (const std::string& n : written) {
      lines.push_back("    double w_" + n + " = c(" + p_ + "_" + n + ", r);");
      scope[n] = "w_" + n;
    }
    stmts(body, scope, lines, "    ", nullptr, {"breakpoint"});
    lines.push_back("    double sum = 0.0;");
    auto value_of = [&](const std::string& n) {
      auto it = scope.find(n);
      if (it == scope.end()) err("current " + n + " is an array field");
      return it->second;
    };
    for (const std::string& n : currents) lines.push_back("    sum = (sum + " + value_of(n) + ");");
    for (const std::string& n : m_.electrode) lines.push_back("    sum = (sum - " + value_of(n) + ");");
    lines.push_back("    if (probe == 0) {");
    lines.push_back("      shifted = sum;");
    if (!ion_currents.empty()) {
      lines.push_back("      if (ish) {");
      for (size_t k = 0; k < ion_currents.size(); ++k)
        lines.push_back("        ish[" + std::to_string(k) + "] = " + value_of(ion_currents[k]) + ";");
      lines.push_back("      }");
    }
    lines.push_back("    } else {");
    lines.push_back("      value = sum;");
    if (!ion_currents.empty()) {
      lines.push_back("      if (ish) {");
      for (size_t k = 0; k < ion_currents.size(); ++k)
        lines.push_back("        ish[" + std::to_string(kIshSlots + k) + "] = " + value_of(ion_currents[k]) + ";");
      lines.push_back("      }");
    }
    std::vector<std::string> stores;
    if (!written.empty()) {
      lines.push_back("      if (store) {");
      for (const std::string& n : written) stores.push_back("        c(" + p_ + "_" + n + ", r) = w_" + n + ";");
      lines.insert(lines.end(), stores.begin(), stores.end());
      lines.push_back("      }");
    }
    lines.push_back("    }");
    lines.push_back("  }");
    auto is_store = [&](const std::string& l) { return in(stores, l); };
    // every column read of the evaluation (outside the store lines), in field order
    std::string text;
    for (size_t k = 0; k < lines.size(); ++k)
      if (!is_store(lines[k])) text += (text.empty() && k == 0 ? "" : "\n") + lines[k];
    std::vector<std::string> reads;
    for (const std::string& n : fields_)
      if (text.find("c(" + p_ + "_" + n + ", r)") != std::string::npos) reads.push_back(n);
    out.push_back("struct " + p_ + "_cur_in {");
    for (const std::string& n : reads) out.push_back("  double " + n + ";");
    if (reads.empty()) out.push_back("  double unused;");
    out.push_back("};");
    out.push_back("MIND_HD " + p_ + "_cur_in " + p_ + "_current_load(const Cols& c, int64_t r) {");
    out.push_back("  " + p_ + "_cur_in in;");
    for (const std::string& n : reads) out.push_back("  in." + n + " = c(" + p_ + "_" + n + ", r);");
    if (reads.empty()) {
      out.push_back("  in.unused = 0.0;");
      out.push_back("  (void)c;");
      out.push_back("  (void)r;");
    }
    out.push_back("  return in;");
    out.push_back("}");
    // replaces one read column's value in <P>_cur_in (a USEION read loaded from
    // the ion row after the column reads were issued)
    out.push_back("MIND_HD void " + p_ + "_cur_set(" + p_ + "_cur_in& in, int f, double x) {");
    out.push_back("  (void)in;");
    out.push_back("  (void)x;");
    out.push_back("  switch (f) {");
    for (const std::string& n : reads) out.push_back("    case " + p_ + "_" + n + ": in." + n + " = x; break;");
    out.push_back("    default: break;");
    out.push_back("  }");
    out.push_back("}");
    out.push_back("MIND_HD void " + p_ + "_current_eval(const Cols& c, int64_t r, const " + p_ +
                  "_cur_in& in, double v, double t, double dt, double& value, double& shifted, bool store, "
                  "double celsius, double* ish) {");
    out.push_back("  (void)in;");
    out.push_back("  (void)ish;");
    for (std::string l : lines) {
      if (!is_store(l)) {
        for (const std::string& n : reads) {
          const std::string from = "c(" + p_ + "_" + n + ", r)", to = "in." + n;
          for (size_t at = l.find(from); at != std::string::npos; at = l.find(from, at + to.size()))
            l.replace(at, from.size(), to);
        }
      }
      out.push_back(l);
    }
    out.push_back("}");
    out.push_back("MIND_HD void " + p_ + "_current(const Cols& c, int64_t r, double v, double t, double dt, "
                  "double& value, double& shifted, bool store, double celsius, double* ish) {");
    out.push_back("  " + p_ + "_current_eval(c, r, " + p_ +
                  "_current_load(c, r), v, t, dt, value, shifted, store, celsius, ish);");
    out.push_back("}");
  }

  // MIND-role mechanisms: BREAKPOINT without its SOLVE statements, run once on the
  // fields (no voltage probes, no membrane current).
  void evaluate(std::vector<std::string>& out) {
    out.push_back("MIND_HD void " + p_ + "_evaluate(const Cols& c, int64_t r, double t, double dt) {");
    out.push_back("  GenCtx x = gen_ctx(0.0, t, dt);");
    out.push_back("  (void)x;");
    const Block* bp = m_.block("breakpoint", "breakpoint");
    if (bp) {
      std::vector<S> body;
      for (const S& s : bp->body)
        if (s->kind != Stmt::SOLVE) body.push_back(s);
      stmts(body, base_scope(), out, "  ");
    }
    out.push_back("}");
  }

  void advance(std::vector<std::string>& out) {
    out.push_back("MIND_HD void " + p_ + "_advance(const Cols& c, int64_t r, double v, double t, double dt, "
                  "double celsius) {");
    out.push_back("  GenCtx x = gen_ctx(v, t, dt, celsius);");
    out.push_back("  (void)x;");
    const Block* bp = m_.block("breakpoint", "breakpoint");
    if (bp) {
      for (const S& st : bp->body) {
        if (st->kind != Stmt::SOLVE) continue;
        if (st->method == "after_cvode" && !proc(st->target))
          err("after_cvode in fixed-step execution requires a PROCEDURE", st->line);
        if (const Block* d = deriv(st->target)) {
          out.push_back("  {");
          if (st->method == "euler")
            euler(*d, out);
          else
            stmts(d->body, base_scope(), out, "    ", "cnexp");
          out.push_back("  }");
        } else if (proc(st->target)) {
          out.push_back("  " + p_ + "_" + st->target + "(c, r, x);");
        } else {
          err("SOLVE of unknown block " + st->target, st->line);
        }
      }
    }
    out.push_back("}");
  }

  // all rates first, then the updates (codegen.jl:663-697)
  void euler(const Block& d, std::vector<std::string>& out) {
    Scope scope = base_scope();
    std::vector<std::pair<std::string, std::string>> rates;
    std::vector<S> body;
    for (const S& st : d.body) {
      if (st->kind == Stmt::ASSIGN && st->derivative) {
        const std::string rn = "d_" + st->target;
        body.push_back(assign_stmt(rn, st->value, false, st->line));
        rates.push_back({st->target, rn});
        scope[rn] = rn;
      } else {
        body.push_back(st);
      }
    }
    for (const auto& r : rates) out.push_back("    double " + r.second + " = 0.0;");
    stmts(body, scope, out, "    ");
    for (const auto& r : rates) {
      auto it = scope.find(r.first);
      if (it == scope.end()) err("unknown name " + r.first);
      const std::string s = it->second;
      out.push_back("    " + s + " = (" + s + " + (x.dt * " + r.second + "));");
    }
  }

  void initial(std::vector<std::string>& out) {
    out.push_back("MIND_HD NetStimOut " + p_ + "_initial(const Cols& c, int64_t r, double v, double celsius, "
                  "Stream* s, uint64_t seed) {");
    out.push_back("  GenCtx x = gen_ctx(v, 0.0, 0.0, celsius, s, seed);");
    out.push_back("  (void)x;");
    if (const Block* b = m_.block("initial", "initial")) stmts(b->body, base_scope(), out, "  ");
    out.push_back("  return x.out;");
    out.push_back("}");
  }

  void receive(std::vector<std::string>& out) {
    out.push_back("MIND_HD NetStimOut " + p_ + "_receive(const Cols& c, int64_t r, double t, double weight, "
                  "double flag, Stream* s, uint64_t seed) {");
    out.push_back("  GenCtx x = gen_ctx(0.0, t, 0.0, 0.0, s, seed);");
    out.push_back("  (void)x;");
    out.push_back("  (void)flag;");
    out.push_back("  (void)weight;");
    if (const Block* b = m_.block("net_receive", "net_receive")) {
      Scope scope = base_scope();
      scope["flag"] = "flag";
      if (b->args.size() > 1) err("NET_RECEIVE takes the connection weight only", b->line);
      for (const std::string& a : b->args) scope[a] = "weight";
      stmts(b->body, scope, out, "  ");
    }
    out.push_back("  return x.out;");
    out.push_back("}");
  }
};

const char* kGenHelpers = R"(struct GenCtx {
  double v, t, dt, celsius;
  Stream* s;
  uint64_t seed;
  NetStimOut out;
};

// Element of an array field addressed by a computed MOD index (0-based), kept
// within the array.
MIND_HD int gen_index(double j, int n) {
  const int k = (int)j;
  return k < 0 ? 0 : (k >= n ? n - 1 : k);
}

MIND_HD GenCtx gen_ctx(double v, double t, double dt = 0.0, double celsius = 0.0, Stream* s = nullptr,
                       uint64_t seed = 0) {
  GenCtx x;
  x.v = v;
  x.t = t;
  x.dt = dt;
  x.celsius = celsius;
  x.s = s;
  x.seed = seed;
  x.out = NetStimOut{0, 0.0, 0, 0.0, 0.0};
  return x;
}

// Nernst potential (mV) and GHK current factor of the nernst / ghk intrinsics
// (src/model/micro/utils/biophysics.jl:43-70), the operations in the
// reference's order.
MIND_HD double gen_nernst(double ci, double co, double z, double celsius) {
  return mind_nernst(ci, co, z, celsius);
}
MIND_HD double gen_ghk(double v, double ci, double co, double z, double celsius) {
  return mind_ghk(v, ci, co, z, celsius);
}
)";

// Builtin mechanisms and their runtime slots (include/mind.h).
const std::vector<std::pair<std::string, std::string>>& builtin_ids() {
  static const std::vector<std::pair<std::string, std::string>> ids = {
      {"hh", "MIND_MECH_HH"},         {"pas", "MIND_MECH_PAS"},
      {"ExpSyn", "MIND_MECH_EXPSYN"}, {"Exp2Syn", "MIND_MECH_EXP2SYN"},
      {"IClamp", "MIND_MECH_ICLAMP"}, {"NetStim", "MIND_MECH_NETSTIM"},
      {"AlphaSynapse", "MIND_MECH_ALPHASYNAPSE"}, {"SEClamp", "MIND_MECH_SECLAMP"},
      {"VClamp", "MIND_MECH_VCLAMP"}};
  return ids;
}

std::string upper(std::string s) {
  for (char& ch : s)
    if (ch >= 'a' && ch <= 'z') ch = (char)(ch - 'a' + 'A');
  return s;
}

// Phase dispatch over the runtime's builtin slots (MIND_MECH_*): a header meant
// for libmind must define every builtin with the runtime's field layout.
std::string dispatch(const std::vector<Mechanism>& mechs) {
  std::map<std::string, const Mechanism*> have;
  for (const Mechanism& m : mechs) have[m.name] = &m;
  bool missing = false;
  for (const auto& b : builtin_ids()) missing |= !have.count(b.first);
  std::vector<std::string> srcs;
  for (const Mechanism& m : mechs) srcs.push_back(basename_of(m.filename) + " (sha256 " + hex16(m.digest()) + ")");
  std::vector<std::string> out;
  out.push_back("constexpr const char* kSource = \"NMODL lowering of " + join(srcs, ", ") + "\";");
  out.push_back(std::string("constexpr bool kComplete = ") + (missing ? "false" : "true") + ";");
  if (missing) return join(out, "\n");
  std::vector<std::string> checks;
  for (const auto& b : builtin_ids())
    for (const std::string& f : have[b.first]->fields())
      checks.push_back(b.first + "_" + f + " == " + b.second.substr(strlen("MIND_MECH_")) + "_" + upper(f));
  out.push_back("static_assert(" + join(checks, " && ") + ", \"generated field layout differs from the runtime's\");");
  out.push_back("constexpr uint64_t kNetStimDigest = NetStim_digest;");
  auto cases = [&](const std::string& pre, const std::string& post, bool plain_only) {
    std::vector<std::string> c;
    for (const auto& b : builtin_ids()) {
      if (plain_only && !have[b.first]->randoms.empty()) continue;
      c.push_back("    case " + b.second + ": " + b.first + pre + post + "; break;");
    }
    return join(c, "\n");
  };
  out.push_back(
      "MIND_HD void current(int mech, const Cols& c, int64_t r, double v, double t, double dt, double& value, "
      "double& shifted, bool store, double celsius, double* ish) {\n  switch (mech) {\n" +
      cases("_current", "(c, r, v, t, dt, value, shifted, store, celsius, ish)", false) +
      "\n    default: value = shifted = 0.0;\n  }\n}");
  out.push_back("template <int MECH>\nstruct CurOf;");
  for (const auto& b : builtin_ids()) {
    const std::string& n = b.first;
    out.push_back("template <>\nstruct CurOf<" + b.second + "> {\n  using In = " + n +
                  "_cur_in;\n  static MIND_HD In load(const Cols& c, int64_t r) { return " + n +
                  "_current_load(c, r); }\n  static MIND_HD void eval(const Cols& c, int64_t r, const In& in, double "
                  "v, double t, double dt, double& value, double& shifted, bool store, double celsius) {\n    " +
                  n + "_current_eval(c, r, in, v, t, dt, value, shifted, store, celsius, nullptr);\n  }\n" +
                  "  static MIND_HD void eval_ish(const Cols& c, int64_t r, const In& in, double v, double t, double "
                  "dt, double& value, double& shifted, bool store, double celsius, double* ish) {\n    " +
                  n + "_current_eval(c, r, in, v, t, dt, value, shifted, store, celsius, ish);\n  }\n" +
                  "  static MIND_HD void set(In& in, int f, double x) { " + n + "_cur_set(in, f, x); }\n};");
  }
  out.push_back("MIND_HD void advance(int mech, const Cols& c, int64_t r, double v, double t, double dt, "
                "double celsius) {\n  switch (mech) {\n" +
                cases("_advance", "(c, r, v, t, dt, celsius)", false) + "\n    default: break;\n  }\n}");
  // membrane mechanisms only: mechanisms with RANDOM streams (NetStim) run their
  // INITIAL and NET_RECEIVE through the event queue with the row's stream
  out.push_back("MIND_HD void initial(int mech, const Cols& c, int64_t r, double v, double celsius) {\n  switch "
                "(mech) {\n" +
                cases("_initial", "(c, r, v, celsius, nullptr, 0)", true) + "\n    default: break;\n  }\n}");
  out.push_back("MIND_HD void receive(int mech, const Cols& c, int64_t r, double weight) {\n  switch (mech) {\n" +
                cases("_receive", "(c, r, 0.0, weight, 0.0, nullptr, 0)", true) + "\n    default: break;\n  }\n}");
  return join(out, "\n");
}

// Every CALL statement named `name` in a statement list (nested blocks included).
bool calls(const std::vector<S>& body, const char* name) {
  for (const S& s : body) {
    if (s->kind == Stmt::CALL && s->value->id == name) return true;
    if (calls(s->body, name) || calls(s->other, name)) return true;
  }
  return false;
}

}  // namespace

std::string builtin_header(const std::vector<std::pair<std::string, std::string>>& sources) {
  std::vector<Mechanism> mechs;
  for (const auto& s : sources) mechs.push_back(parse(s.second, s.first));
  std::set<std::string> names;
  for (const Mechanism& m : mechs)
    if (!names.insert(m.name).second) raise_at("<input>", 1, "duplicate mechanism names");
  std::vector<std::string> parts;
  parts.push_back(std::string("// Generated by mind_amd.nmodl from MOD sources; do not edit.\n#pragma once\n\n"
                              "#include \"mech.h\"\n\nnamespace mind {\nnamespace gen {\n\n") +
                  kGenHelpers);
  for (const Mechanism& m : mechs) parts.push_back(Gen(m, m.name).emit());
  parts.push_back(dispatch(mechs));
  parts.push_back("\n}  // namespace gen\n}  // namespace mind\n");
  return join(parts, "\n");
}

std::string user_prelude() { return kGenHelpers; }

std::string macro_header(const std::vector<std::pair<std::string, std::string>>& sources) {
  std::vector<Mechanism> mechs;
  for (const auto& s : sources) mechs.push_back(parse(s.second, s.first));
  std::set<std::string> names;
  std::vector<std::string> srcs;
  for (const Mechanism& m : mechs) {
    if (m.role.empty()) raise_at(m.filename, 1, "no MIND ROLE: " + m.name + " is not a macro-step mechanism");
    if (!names.insert(m.name).second) raise_at("<input>", 1, "duplicate mechanism names");
    srcs.push_back(basename_of(m.filename) + " (" + m.role + ", sha256 " + hex16(m.digest()) + ")");
  }
  std::vector<std::string> parts;
  parts.push_back(std::string("// Generated by nmodlc from MIND-role MOD sources; do not edit.\n#pragma once\n\n"
                              "#include \"mech.h\"\n\nnamespace mind {\nnamespace gen_macro {\n\n") +
                  kGenHelpers);
  for (const Mechanism& m : mechs) parts.push_back(Gen(m, m.name).emit());
  parts.push_back("constexpr const char* kSource = \"NMODL lowering of " + join(srcs, ", ") + "\";");
  parts.push_back("\n}  // namespace gen_macro\n}  // namespace mind\n");
  return join(parts, "\n");
}

UserMechanism lower_user(const std::string& filename, const std::string& text, const std::string& prefix) {
  const Mechanism m = parse(text, filename);
  auto refuse = [&](const std::string& msg) { raise_at(filename, 1, msg); };
  if (m.kind == "artificial_cell")
    refuse("ARTIFICIAL_CELL " + m.name + ": runtime mechanisms are membrane mechanisms (SUFFIX or POINT_PROCESS)");
  if (!m.role.empty())
    refuse("MIND ROLE " + m.role + ": the macro step's role mechanisms are lowered into its kernels at library build");
  if (!m.randoms.empty()) refuse("RANDOM variables are supported for the builtin NetStim only");
  for (const Block& b : m.blocks)
    if (calls(b.body, "net_send") || calls(b.body, "net_event"))
      raise_at(filename, b.line, "net_send and net_event are supported for the builtin NetStim only");
  struct IonDefaults { const char* ion; double e, in, out; };
  // default_ions (src/model/micro/utils/biophysics.jl:31-42); a species the
  // model declares itself starts at 1 mM on both sides (model.jl:696-703)
  static const IonDefaults kIons[] = {{"na", 50.0, 10.0, 140.0},
                                      {"k", -77.0, 54.4, 2.5},
                                      {"ca", 132.4579341637009, 5.0e-5, 2.0}};
  Gen g(m, prefix);
  std::map<std::string, double> ion_value;
  std::vector<IonUse> uses;
  for (const Ion& ion : m.ions) {
    const IonDefaults* d = nullptr;
    for (const IonDefaults& k : kIons)
      if (ion.ion == k.ion) d = &k;
    const double start[4] = {d ? d->e : 0.0, d ? d->in : 1.0, d ? d->out : 1.0, 0.0};
    // ion columns: reversal, internal and external concentration, current
    const std::string names[4] = {"e" + ion.ion, ion.ion + "i", ion.ion + "o", "i" + ion.ion};
    auto slot_of = [&](const std::string& n) {
      for (int k = 0; k < 4; ++k)
        if (n == names[k]) return k;
      return -1;
    };
    auto field_of = [&](const std::string& n, const char* access) {
      bool scalar = false;
      for (const std::string& f : g.fields()) scalar |= f == n && !m.extents.count(n);
      if (!scalar) refuse("USEION " + ion.ion + " " + access + " " + n + ": not a declared scalar variable");
      return g.column(n);
    };
    IonUse u;
    u.name = ion.ion;
    u.has_valence = ion.has_valence;
    u.valence = ion.valence;
    for (const std::string& w : ion.writes) {
      const int k = slot_of(w);
      if (k < 0) refuse("USEION " + ion.ion + " WRITE " + w + ": not an ion variable of " + ion.ion);
      u.write[k] = field_of(w, "WRITE");
    }
    // written concentrations are loaded too (utils/build.jl:118-124)
    std::vector<std::string> reads = ion.reads;
    for (int k : {1, 2})
      if (u.write[k] >= 0 && std::find(reads.begin(), reads.end(), names[k]) == reads.end()) reads.push_back(names[k]);
    for (const std::string& r : reads) {
      const int k = slot_of(r);
      if (k < 0) refuse("USEION " + ion.ion + " READ " + r + ": not an ion variable of " + ion.ion);
      u.read[k] = field_of(r, "READ");
      ion_value[r] = start[k];
    }
    uses.push_back(u);
  }
  // A GLOBAL is one value every instance shares (the reference's shared storage,
  // initialize.jl:97-98; global_ref, references.jl:147-154). Instances that assign
  // one run one after another there (serial instances, initialize.jl:959-1086); the
  // device phases run instances side by side, so a GLOBAL may be read, not assigned.
  std::set<std::string> globals;
  for (const std::string& n : m.globals)
    if (std::count(m.ranges.begin(), m.ranges.end(), n) == 1) globals.insert(n);  // not also RANGE
  for (const Block& b : m.blocks)
    for (const std::string& n : g.block_writes(b))
      if (globals.count(n))
        raise_at(filename, b.line,
                 "GLOBAL " + n + " is assigned: instances sharing a GLOBAL would run one after another; "
                 "declare it RANGE");
  UserMechanism u;
  u.name = m.name;
  u.filename = filename;
  u.kind = m.kind == "point_process" ? 1 : 0;
  u.digest = m.digest();
  u.code = g.emit();
  u.ions = uses;
  std::map<std::string, double> defaults;
  for (const Decl& d : m.decls)
    if (d.storage == "parameter" && d.has_default) defaults[d.name] = d.value;
  std::set<std::string> states;
  for (const Decl& d : m.decls)
    if (d.storage == "state") states.insert(d.name);
  for (const std::string& n : g.fields()) {
    const int ext = m.extents.count(n) ? m.extents.at(n) : 0;
    for (int k = 0; k < std::max(ext, 1); ++k) {
      u.columns.push_back(ext ? n + "[" + std::to_string(k + 1) + "]" : n);
      double v = 0.0;
      if (k == 0 && defaults.count(n)) v = defaults[n];
      if (ion_value.count(n)) v = ion_value[n];
      u.defaults.push_back(v);
      u.state_column.push_back(states.count(n) ? 1 : 0);
      u.global_column.push_back(globals.count(n) ? 1 : 0);
    }
  }
  // the fields NET_RECEIVE touches, from its lowered text
  {
    const std::string head = "MIND_HD NetStimOut " + prefix + "_receive(";
    const size_t at = u.code.find(head);
    const std::string body = at == std::string::npos ? "" : u.code.substr(at, u.code.find("\n}", at) - at);
    if (body.find("x.celsius") != std::string::npos)
      refuse("celsius in NET_RECEIVE: a runtime mechanism's NET_RECEIVE sees the connection weight only");
    for (const std::string& n : g.fields())
      if (body.find("c(" + prefix + "_" + n + ", r)") != std::string::npos) u.receive_fields.push_back(g.column(n));
  }
  u.has_net_receive = m.block("net_receive", "net_receive") != nullptr;
  const Block* bp = m.block("breakpoint", "breakpoint");
  if (bp)
    for (const S& s : bp->body) u.has_state |= s->kind == Stmt::SOLVE;
  return u;
}

}  // namespace nmodl
}  // namespace mind
