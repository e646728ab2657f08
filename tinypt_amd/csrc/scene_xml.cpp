// scene_xml.cpp -- scene files on the host.
//
//   * tpt_scene_load_xml: the Mitsuba-XML subset of load_scene (mitsuba.rs:137-595) -- its
//     element mapping, defaults, warnings and fallbacks -- turned into a tpt_scene_desc whose
//     arrays the returned handle owns.  The XML reader is a small one of this file: elements,
//     attributes (single or double quotes, the five named entities and numeric references),
//     comments, processing instructions and DOCTYPE are understood; text content is ignored,
//     as the reference's reader trims and never reads it.
//   * tpt_obj_load: load_obj_triangles (obj_loader.rs:101-109).
//   * load_hdr: read_hdr (hdr.rs:20-62) for envmap emitters.  OpenEXR maps go through a
//     third-party decoder in the reference (exr.rs:20-41) and are not read here: such an
//     envmap is skipped with a warning, which is what the reference does on a failed load.
//
// Compiled with -ffp-contract=off: transforms, colours and camera set-up restate the
// reference's f64 arithmetic operation by operation.
#include "scene_xml.h"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

namespace tpt {
namespace {

const double kPi = 3.14159265358979323846;
const double kTau = 6.28318530717958647692;

// ------------------------------------------------------------------ XML tree
struct Element {
  std::string tag;
  std::vector<std::pair<std::string, std::string>> attrs;
  std::vector<Element> children;

  const std::string* attr(const char* key) const {
    for (const auto& a : attrs)
      if (a.first == key) return &a.second;
    return nullptr;
  }
  std::string typ() const {
    const std::string* t = attr("type");
    return t ? *t : std::string();
  }
  // first child with this tag whose name attribute equals `name` (mitsuba.rs:59-64)
  const Element* prop(const char* tag_, const char* name) const {
    for (const Element& c : children) {
      const std::string* n = c.attr("name");
      if (c.tag == tag_ && n && *n == name) return &c;
    }
    return nullptr;
  }
  const Element* child_tag(const char* tag_) const {
    for (const Element& c : children)
      if (c.tag == tag_) return &c;
    return nullptr;
  }
};

struct XmlReader {
  const std::string& s;
  size_t i = 0;
  std::string err;
  explicit XmlReader(const std::string& src) : s(src) {}

  bool fail(const std::string& m) {
    if (err.empty()) err = m + " at byte " + std::to_string(i);
    return false;
  }
  bool starts(const char* p) const { return s.compare(i, std::strlen(p), p) == 0; }
  void skip_ws() {
    while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) i++;
  }
  bool skip_past(const char* p) {
    size_t j = s.find(p, i);
    if (j == std::string::npos) return fail(std::string("unterminated construct, missing '") + p + "'");
    i = j + std::strlen(p);
    return true;
  }
  static bool name_char(char c) {
    return std::isalnum((unsigned char)c) || c == '_' || c == '-' || c == '.' || c == ':' || (unsigned char)c >= 0x80;
  }
  bool read_name(std::string* out) {
    size_t j = i;
    while (j < s.size() && name_char(s[j])) j++;
    if (j == i) return fail("expected a name");
    *out = s.substr(i, j - i);
    i = j;
    return true;
  }
  static void put_utf8(std::string* o, unsigned long cp) {
    if (cp < 0x80) {
      o->push_back((char)cp);
    } else if (cp < 0x800) {
      o->push_back((char)(0xC0 | (cp >> 6)));
      o->push_back((char)(0x80 | (cp & 0x3F)));
    } else if (cp < 0x10000) {
      o->push_back((char)(0xE0 | (cp >> 12)));
      o->push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
      o->push_back((char)(0x80 | (cp & 0x3F)));
    } else {
      o->push_back((char)(0xF0 | (cp >> 18)));
      o->push_back((char)(0x80 | ((cp >> 12) & 0x3F)));
      o->push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
      o->push_back((char)(0x80 | (cp & 0x3F)));
    }
  }
  bool unescape(const std::string& raw, std::string* out) {
    out->clear();
    for (size_t k = 0; k < raw.size(); k++) {
      if (raw[k] != '&') {
        out->push_back(raw[k]);
        continue;
      }
      size_t e = raw.find(';', k);
      if (e == std::string::npos) return fail("unterminated entity reference");
      std::string ent = raw.substr(k + 1, e - k - 1);
      if (ent == "lt") out->push_back('<');
      else if (ent == "gt") out->push_back('>');
      else if (ent == "amp") out->push_back('&');
      else if (ent == "quot") out->push_back('"');
      else if (ent == "apos") out->push_back('\'');
      else if (ent.size() > 1 && ent[0] == '#') {
        char* end = nullptr;
        unsigned long cp = ent[1] == 'x' ? std::strtoul(ent.c_str() + 2, &end, 16) : std::strtoul(ent.c_str() + 1, &end, 10);
        if (!end || *end != '\0' || cp > 0x10FFFF) return fail("bad character reference");
        put_utf8(out, cp);
      } else {
        return fail("unknown entity &" + ent + ";");
      }
      k = e;
    }
    return true;
  }
  // <tag attr="v" ...> or <tag .../>; `empty` tells which
  bool read_start(Element* el, bool* empty) {
    i++;  // '<'
    if (!read_name(&el->tag)) return false;
    for (;;) {
      skip_ws();
      if (i >= s.size()) return fail("unterminated start tag");
      if (s[i] == '/') {
        if (!starts("/>")) return fail("expected '/>'");
        i += 2;
        *empty = true;
        return true;
      }
      if (s[i] == '>') {
        i++;
        *empty = false;
        return true;
      }
      std::string key, val;
      if (!read_name(&key)) return false;
      skip_ws();
      if (i >= s.size() || s[i] != '=') return fail("expected '=' after attribute " + key);
      i++;
      skip_ws();
      if (i >= s.size() || (s[i] != '"' && s[i] != '\'')) return fail("expected a quoted attribute value");
      char q = s[i++];
      size_t e = s.find(q, i);
      if (e == std::string::npos) return fail("unterminated attribute value");
      if (!unescape(s.substr(i, e - i), &val)) return false;
      i = e + 1;
      for (const auto& a : el->attrs)
        if (a.first == key) return fail("duplicated attribute " + key);
      el->attrs.emplace_back(key, val);
    }
  }
  bool parse(Element* root) {
    std::vector<Element> stack;
    bool have_root = false;
    auto attach = [&](Element&& el) -> bool {
      if (!stack.empty()) {
        stack.back().children.push_back(std::move(el));
      } else {
        *root = std::move(el);  // a later top-level element replaces an earlier one (mitsuba.rs:283-289)
        have_root = true;
      }
      return true;
    };
    while (i < s.size()) {
      size_t lt = s.find('<', i);
      if (lt == std::string::npos) break;  // trailing text
      i = lt;
      if (starts("<!--")) {
        if (!skip_past("-->")) return false;
      } else if (starts("<?")) {
        if (!skip_past("?>")) return false;
      } else if (starts("<![CDATA[")) {
        if (!skip_past("]]>")) return false;
      } else if (starts("<!")) {
        if (!skip_past(">")) return false;
      } else if (starts("</")) {
        i += 2;
        std::string name;
        if (!read_name(&name)) return false;
        skip_ws();
        if (i >= s.size() || s[i] != '>') return fail("expected '>' in end tag");
        i++;
        if (stack.empty() || stack.back().tag != name) return fail("unbalanced XML (</" + name + ">)");
        Element el = std::move(stack.back());
        stack.pop_back();
        if (!attach(std::move(el))) return false;
      } else {
        Element el;
        bool empty = false;
        if (!read_start(&el, &empty)) return false;
        if (empty) {
          if (!attach(std::move(el))) return false;
        } else {
          stack.push_back(std::move(el));
        }
      }
    }
    if (!stack.empty()) return fail("unbalanced XML (<" + stack.back().tag + "> not closed)");
    if (!have_root) return fail("empty XML");
    return true;
  }
};

// ------------------------------------------------------------------ values (mitsuba.rs:70-122)
std::string trim(const std::string& x) {
  size_t a = 0, b = x.size();
  while (a < b && std::isspace((unsigned char)x[a])) a++;
  while (b > a && std::isspace((unsigned char)x[b - 1])) b--;
  return x.substr(a, b - a);
}

// f64::from_str on a trimmed string: decimal forms, "inf" / "infinity" / "nan" (any case);
// no hexadecimal floats.
bool parse_f64(const std::string& raw, double* out) {
  std::string t = trim(raw);
  if (t.empty()) return false;
  std::string body = (t[0] == '+' || t[0] == '-') ? t.substr(1) : t;
  std::string low = body;
  for (char& c : low) c = (char)std::tolower((unsigned char)c);
  if (low == "inf" || low == "infinity" || low == "nan") {
    double v = low == "nan" ? NAN : INFINITY;
    *out = t[0] == '-' ? -v : v;
    return true;
  }
  for (char c : body)
    if (!(std::isdigit((unsigned char)c) || c == '.' || c == 'e' || c == 'E' || c == '+' || c == '-')) return false;
  errno = 0;
  char* end = nullptr;
  double v = std::strtod(t.c_str(), &end);
  if (!end || *end != '\0') return false;
  *out = v;
  return true;
}

// usize::from_str on a trimmed string
bool parse_usize(const std::string& raw, uint64_t* out) {
  std::string t = trim(raw);
  if (!t.empty() && t[0] == '+') t = t.substr(1);
  if (t.empty() || t.size() > 19) return false;
  uint64_t v = 0;
  for (char c : t) {
    if (!std::isdigit((unsigned char)c)) return false;
    v = v * 10 + (uint64_t)(c - '0');
  }
  *out = v;
  return true;
}

struct V3 {
  double x, y, z;
};

// "x,y,z" / "x y z" / one broadcast scalar (mitsuba.rs:111-122)
bool parse_vec3(const std::string& s, V3* out) {
  std::vector<double> parts;
  std::string tok;
  auto flush = [&]() -> bool {
    if (!trim(tok).empty()) {
      double v;
      if (!parse_f64(tok, &v)) return false;
      parts.push_back(v);
    }
    tok.clear();
    return true;
  };
  for (char c : s) {
    if (c == ',' || c == ' ') {
      if (!flush()) return false;
    } else {
      tok.push_back(c);
    }
  }
  if (!flush()) return false;
  if (parts.size() == 3) *out = V3{parts[0], parts[1], parts[2]};
  else if (parts.size() == 1) *out = V3{parts[0], parts[0], parts[0]};
  else return false;
  return true;
}

bool el_float(const Element& e, const char* name, double* v) {
  const Element* p = e.prop("float", name);
  const std::string* a = p ? p->attr("value") : nullptr;
  return a && parse_f64(*a, v);
}
bool el_int(const Element& e, const char* name, uint64_t* v) {
  const Element* p = e.prop("integer", name);
  const std::string* a = p ? p->attr("value") : nullptr;
  return a && parse_usize(*a, v);
}
const std::string* el_string(const Element& e, const char* name) {
  const Element* p = e.prop("string", name);
  return p ? p->attr("value") : nullptr;
}
bool el_point(const Element& e, const char* name, V3* out) {
  const Element* p = e.prop("point", name);
  if (!p) return false;
  const std::string *x = p->attr("x"), *y = p->attr("y"), *z = p->attr("z");
  if (x && y && z) return parse_f64(*x, &out->x) && parse_f64(*y, &out->y) && parse_f64(*z, &out->z);
  const std::string* v = p->attr("value");
  return v && parse_vec3(*v, out);
}
double srgb_to_linear(double c) {  // math.rs:176-182
  return c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4);
}
// <rgb> is linear, <srgb> gamma-decoded (mitsuba.rs:93-104)
bool el_color(const Element& e, const char* name, V3* out) {
  if (const Element* p = e.prop("rgb", name)) {
    const std::string* v = p->attr("value");
    return v && parse_vec3(*v, out);
  }
  if (const Element* p = e.prop("srgb", name)) {
    const std::string* v = p->attr("value");
    if (!v || !parse_vec3(*v, out)) return false;
    *out = V3{srgb_to_linear(out->x), srgb_to_linear(out->y), srgb_to_linear(out->z)};
    return true;
  }
  return false;
}

// ------------------------------------------------------------------ transforms (transform.rs:22-112)
struct Affine {
  double a[9];
  double t[3];
};
Affine identity() { return Affine{{1, 0, 0, 0, 1, 0, 0, 0, 1}, {0, 0, 0}}; }
void mat_mul(const double* x, const double* y, double* r) {  // Mat3::mul
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) r[3 * i + j] = x[3 * i] * y[j] + x[3 * i + 1] * y[3 + j] + x[3 * i + 2] * y[6 + j];
}
void mat_vec(const double* m, const double* v, double* r) {  // Mat3::mul_vec
  for (int i = 0; i < 3; i++) r[i] = m[3 * i] * v[0] + m[3 * i + 1] * v[1] + m[3 * i + 2] * v[2];
}
// self . inner (transform.rs:70-72)
Affine compose(const Affine& self, const Affine& inner) {
  Affine r;
  mat_mul(self.a, inner.a, r.a);
  double at[3];
  mat_vec(self.a, inner.t, at);
  for (int k = 0; k < 3; k++) r.t[k] = at[k] + self.t[k];
  return r;
}
Affine rotation(V3 axis, double deg) {  // Rodrigues (transform.rs:101-112)
  double len = std::sqrt(axis.x * axis.x + axis.y * axis.y + axis.z * axis.z);
  double inv = 1.0 / std::max(len, 1e-30);  // Vec3::norm (math.rs:69-72)
  double x = axis.x * inv, y = axis.y * inv, z = axis.z * inv;
  double rad = deg * (kPi / 180.0);  // f64::to_radians
  double s = std::sin(rad), c = std::cos(rad);
  double one_c = 1.0 - c;
  return Affine{{c + x * x * one_c, x * y * one_c - z * s, x * z * one_c + y * s,  //
                 y * x * one_c + z * s, c + y * y * one_c, y * z * one_c - x * s,  //
                 z * x * one_c - y * s, z * y * one_c + x * s, c + z * z * one_c},
                {0, 0, 0}};
}

// ------------------------------------------------------------------ canonical shapes (mitsuba.rs:414-470)
void push_tri(std::vector<double>* t, V3 a, V3 b, V3 c) {
  for (V3 v : {a, b, c}) {
    t->push_back(v.x);
    t->push_back(v.y);
    t->push_back(v.z);
  }
}
std::vector<double> rectangle_tris() {
  std::vector<double> t;
  V3 a{-1, -1, 0}, b{1, -1, 0}, c{1, 1, 0}, d{-1, 1, 0};
  push_tri(&t, a, b, c);
  push_tri(&t, a, c, d);
  return t;
}
std::vector<double> cube_tris() {
  const V3 v[8] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {-1, 1, 1}};
  const int q[6][4] = {{0, 1, 2, 3}, {4, 7, 6, 5}, {0, 4, 5, 1}, {1, 5, 6, 2}, {2, 6, 7, 3}, {3, 7, 4, 0}};
  std::vector<double> t;
  for (const auto& f : q) {
    push_tri(&t, v[f[0]], v[f[1]], v[f[2]]);
    push_tri(&t, v[f[0]], v[f[2]], v[f[3]]);
  }
  return t;
}
std::vector<double> disk_tris() {
  const int n = 64;
  std::vector<double> t;
  for (int i = 0; i < n; i++) {
    double a0 = kTau * (double)i / (double)n, a1 = kTau * (double)(i + 1) / (double)n;
    push_tri(&t, V3{0, 0, 0}, V3{std::cos(a0), std::sin(a0), 0}, V3{std::cos(a1), std::sin(a1), 0});
  }
  return t;
}

std::string dir_of(const std::string& path) {
  size_t k = path.find_last_of('/');
  return k == std::string::npos ? std::string() : path.substr(0, k);
}
std::string resolve(const std::string& base, const std::string& f) {  // mitsuba.rs:472-481
  if (!f.empty() && f[0] == '/') return f;
  return base.empty() ? f : base + "/" + f;
}

bool read_file(const std::string& path, std::string* out, std::string* err) {
  std::ifstream in(path, std::ios::binary);
  if (!in) {
    *err = "cannot open " + path + ": " + std::strerror(errno);
    return false;
  }
  std::ostringstream ss;
  ss << in.rdbuf();
  *out = ss.str();
  return true;
}

}  // namespace

// ------------------------------------------------------------------ OBJ (obj_loader.rs:25-69)
bool load_obj(const std::string& path, std::vector<double>* tris, std::string* err) {
  std::string text;
  if (!read_file(path, &text, err)) return false;
  std::vector<V3> pos;
  std::istringstream in(text);
  std::string line;
  auto f64_or0 = [](const std::string& tok) {
    double v;
    return parse_f64(tok, &v) ? v : 0.0;
  };
  while (std::getline(in, line)) {
    std::string l = trim(line);
    if (l.empty() || l[0] == '#') continue;
    if (l.compare(0, 2, "v ") == 0) {
      std::istringstream ls(l.substr(2));
      std::string x, y, z;
      ls >> x >> y >> z;
      pos.push_back(V3{x.empty() ? 0.0 : f64_or0(x), y.empty() ? 0.0 : f64_or0(y), z.empty() ? 0.0 : f64_or0(z)});
    } else if (l.compare(0, 2, "f ") == 0) {
      std::istringstream ls(l.substr(2));
      std::string tok;
      std::vector<size_t> face;
      while (ls >> tok) {
        std::string first = tok.substr(0, tok.find('/'));
        // i32::parse, 0 on failure; 1-based, negative = from the end (obj_loader.rs:10-21)
        long long idx = 0;
        {
          std::string t = first;
          bool ok = !t.empty();
          size_t k = (t[0] == '+' || t[0] == '-') ? 1 : 0;
          if (k == t.size()) ok = false;
          for (size_t m = k; ok && m < t.size(); m++) ok = std::isdigit((unsigned char)t[m]) != 0;
          if (ok && t.size() < 12) idx = std::stoll(t);
          if (idx > 2147483647LL || idx < -2147483648LL) idx = 0;
        }
        const long long n = (long long)pos.size();
        if (idx > 0 && idx - 1 < n) face.push_back((size_t)(idx - 1));
        else if (idx < 0 && n + idx >= 0) face.push_back((size_t)(n + idx));
      }
      for (size_t k = 1; face.size() >= 3 && k + 1 < face.size(); k++) push_tri(tris, pos[face[0]], pos[face[k]], pos[face[k + 1]]);
    }
  }
  return true;
}

// ------------------------------------------------------------------ RGBE (hdr.rs:20-62, 90-147, 197-209)
bool load_hdr(const std::string& path, int* w_out, int* h_out, std::vector<double>* rgb, std::string* err) {
  std::string s;
  if (!read_file(path, &s, err)) return false;
  size_t i = 0;
  auto read_line = [&](std::string* l) -> bool {
    if (i >= s.size()) return false;
    size_t e = s.find('\n', i);
    if (e == std::string::npos) e = s.size() - 1;
    *l = s.substr(i, e - i + 1);
    i = e + 1;
    return true;
  };
  std::string line;
  if (!read_line(&line) || (line.compare(0, 10, "#?RADIANCE") != 0 && line.compare(0, 6, "#?RGBE") != 0)) {
    *err = "invalid HDR header";
    return false;
  }
  for (;;) {
    if (!read_line(&line)) {
      *err = "missing resolution";
      return false;
    }
    if (trim(line).empty()) break;
  }
  if (!read_line(&line)) line.clear();
  std::istringstream rs(line);
  std::string p0, p1, p2, p3, extra;
  rs >> p0 >> p1 >> p2 >> p3;
  uint64_t hh = 0, ww = 0;
  if (p3.empty() || (rs >> extra)) {
    *err = "invalid resolution line";
    return false;
  }
  if (!parse_usize(p1, &hh)) { *err = "invalid height"; return false; }
  if (!parse_usize(p3, &ww)) { *err = "invalid width"; return false; }
  const char ys = p0[0], xs = p2[0];
  const size_t w = (size_t)ww, h = (size_t)hh;
  // an empty image has no texel to sample; the product is checked without wrapping (both
  // factors may be up to 19 digits long)
  if (w == 0 || h == 0) { *err = "empty HDR image"; return false; }
  size_t texels = 0;
  if (__builtin_mul_overflow(w, h, &texels) || texels > ((size_t)1 << 28)) { *err = "image too large"; return false; }
  rgb->assign(3 * w * h, 0.0);
  std::vector<uint8_t> scan(4 * w);
  auto need = [&](size_t n) -> bool {
    if (i + n > s.size()) { *err = "unexpected end of HDR data"; return false; }
    return true;
  };
  for (size_t y = 0; y < h; y++) {
    if (w < 8 || w > 0x7fff) {
      if (!need(4 * w)) return false;
      std::memcpy(scan.data(), s.data() + i, 4 * w);
      i += 4 * w;
    } else {
      if (!need(4)) return false;
      const uint8_t* hd = (const uint8_t*)s.data() + i;
      i += 4;
      if (hd[0] != 2 || hd[1] != 2 || (((size_t)hd[2] << 8) | hd[3]) != w) {  // flat scanline
        std::memcpy(scan.data(), hd, 4);
        if (!need(4 * w - 4)) return false;
        std::memcpy(scan.data() + 4, s.data() + i, 4 * w - 4);
        i += 4 * w - 4;
      } else {
        for (int ch = 0; ch < 4; ch++) {
          size_t x = 0;
          while (x < w) {
            if (!need(1)) return false;
            uint8_t code = (uint8_t)s[i++];
            // a run past the end of the scanline would index out of bounds in the reference
            // (a panic there); here it is an error
            if (code > 128) {
              size_t count = code - 128u;
              if (!need(1)) return false;
              uint8_t v = (uint8_t)s[i++];
              if (x + count > w) { *err = "RLE run overflows the scanline"; return false; }
              for (size_t k = 0; k < count; k++) scan[4 * (x++) + (size_t)ch] = v;
            } else {
              size_t count = code;
              if (x + count > w) { *err = "RLE run overflows the scanline"; return false; }
              if (!need(count)) return false;
              for (size_t k = 0; k < count; k++) scan[4 * (x++) + (size_t)ch] = (uint8_t)s[i++];
            }
          }
        }
      }
    }
    for (size_t x = 0; x < w; x++) {
      size_t sx = xs == '-' ? w - 1 - x : x;
      size_t dy = ys == '-' ? y : h - 1 - y;
      const uint8_t* e = &scan[4 * sx];
      double* o = &(*rgb)[3 * (dy * w + x)];
      if (e[3] == 0) {
        o[0] = o[1] = o[2] = 0.0;
      } else {
        double f = std::ldexp(1.0, (int)e[3] - 128 - 8);  // 2^(E - 136), exact
        o[0] = (double)e[0] * f;
        o[1] = (double)e[1] * f;
        o[2] = (double)e[2] * f;
      }
    }
  }
  *w_out = (int)w;
  *h_out = (int)h;
  return true;
}

}  // namespace tpt

// ------------------------------------------------------------------ the loaded scene
struct tpt_loaded_scene {
  tpt_scene_desc desc{};
  std::vector<int32_t> mat_kind;
  std::vector<double> mat_param;
  std::vector<double> sph_c, sph_r;
  std::vector<int32_t> sph_mat;
  std::vector<int64_t> mesh_begin{0};
  std::vector<double> tri_v;
  std::vector<int32_t> tri_mat;
  std::vector<int32_t> inst_mesh;
  std::vector<double> inst_xform;
  std::vector<int32_t> inst_override;
  std::vector<double> env_rgb;
  std::string warnings;

  void warn(const std::string& m) {
    std::fprintf(stderr, "[mitsuba] warning: %s\n", m.c_str());
    warnings += m;
    warnings += '\n';
  }
  int add_material(int kind, double a, double b, double c, double d) {
    mat_kind.push_back(kind);
    for (double v : {a, b, c, d}) mat_param.push_back(v);
    return (int)mat_kind.size() - 1;
  }
  void finish() {
    tpt_scene_desc& d = desc;
    d.n_materials = (int32_t)mat_kind.size();
    d.mat_kind = mat_kind.data();
    d.mat_param = mat_param.data();
    d.n_spheres = (int32_t)sph_mat.size();
    d.sphere_center = sph_c.data();
    d.sphere_radius = sph_r.data();
    d.sphere_mat = sph_mat.data();
    d.n_meshes = (int32_t)(mesh_begin.size() - 1);
    d.mesh_tri_begin = mesh_begin.data();
    d.tri_vertices = tri_v.data();
    d.tri_mat = tri_mat.data();
    d.n_instances = (int32_t)inst_mesh.size();
    d.inst_mesh = inst_mesh.data();
    d.inst_xform = inst_xform.data();
    d.inst_mat_override = inst_override.data();
    d.env_rgb = env_rgb.data();
  }
};

namespace {
using namespace tpt;

// parse_bsdf (mitsuba.rs:558-595): kind and 4 parameters
void parse_bsdf(const Element& el, tpt_loaded_scene* L, int* kind, double p[4]) {
  const std::string t = el.typ();
  V3 c;
  if (t == "twosided") {
    if (const Element* inner = el.child_tag("bsdf")) return parse_bsdf(*inner, L, kind, p);
    *kind = TPT_MAT_LAMBERT;
    p[0] = p[1] = p[2] = 0.5;
    p[3] = 0.0;
  } else if (t == "diffuse") {
    if (!el_color(el, "reflectance", &c)) c = V3{0.5, 0.5, 0.5};
    *kind = TPT_MAT_LAMBERT;
    p[0] = c.x; p[1] = c.y; p[2] = c.z; p[3] = 0.0;
  } else if (t == "conductor") {
    if (!el_color(el, "specular_reflectance", &c)) c = V3{1, 1, 1};
    *kind = TPT_MAT_METAL;
    p[0] = c.x; p[1] = c.y; p[2] = c.z; p[3] = 0.0;
  } else if (t == "roughconductor") {
    const std::string* dist = el_string(el, "distribution");
    if (dist && *dist != "ggx") L->warn("roughconductor distribution '" + *dist + "' unsupported; using ggx");
    if (!el_color(el, "specular_reflectance", &c)) c = V3{1, 1, 1};
    double alpha;
    if (!el_float(el, "alpha", &alpha)) alpha = 0.1;
    *kind = TPT_MAT_GGX;
    p[0] = c.x; p[1] = c.y; p[2] = c.z; p[3] = alpha;
  } else if (t == "dielectric" || t == "thindielectric" || t == "roughdielectric") {
    double int_ior, ext_ior;
    if (!el_float(el, "int_ior", &int_ior)) int_ior = 1.5;
    if (!el_float(el, "ext_ior", &ext_ior)) ext_ior = 1.0;
    if (!el_color(el, "absorption", &c)) c = V3{0, 0, 0};
    *kind = TPT_MAT_DIELECTRIC;
    p[0] = c.x; p[1] = c.y; p[2] = c.z; p[3] = int_ior / ext_ior;
  } else {
    L->warn("unsupported bsdf type '" + t + "'; defaulting to diffuse");
    if (!el_color(el, "reflectance", &c)) c = V3{0.5, 0.5, 0.5};
    *kind = TPT_MAT_LAMBERT;
    p[0] = c.x; p[1] = c.y; p[2] = c.z; p[3] = 0.0;
  }
}

bool transform_op(const Element& el, tpt_loaded_scene* L, Affine* out) {  // mitsuba.rs:493-541
  auto xyz = [&](double def) {
    auto g = [&](const char* k) {
      const std::string* a = el.attr(k);
      double v;
      return a && parse_f64(*a, &v) ? v : def;
    };
    return V3{g("x"), g("y"), g("z")};
  };
  if (el.tag == "translate") {
    V3 t = xyz(0.0);
    *out = identity();
    out->t[0] = t.x; out->t[1] = t.y; out->t[2] = t.z;
    return true;
  }
  if (el.tag == "scale") {
    const std::string* v = el.attr("value");
    double u;
    V3 s = (v && parse_f64(*v, &u)) ? V3{u, u, u} : xyz(1.0);
    *out = Affine{{s.x, 0, 0, 0, s.y, 0, 0, 0, s.z}, {0, 0, 0}};
    return true;
  }
  if (el.tag == "rotate") {
    V3 axis = xyz(0.0);
    const std::string* a = el.attr("angle");
    double angle;
    if (!a || !parse_f64(*a, &angle)) angle = 0.0;
    if (std::sqrt(axis.x * axis.x + axis.y * axis.y + axis.z * axis.z) < 1e-12) {
      L->warn("rotate with zero axis; ignored");
      return false;
    }
    *out = rotation(axis, angle);
    return true;
  }
  if (el.tag == "matrix") {
    const std::string* v = el.attr("value");
    if (!v) return false;
    std::vector<double> vals;
    std::string tok;
    auto flush = [&]() {
      double x;
      if (!trim(tok).empty() && parse_f64(tok, &x)) vals.push_back(x);
      tok.clear();
    };
    for (char c : *v) {
      if (c == ',' || c == ' ') flush();
      else tok.push_back(c);
    }
    flush();
    if (vals.size() != 16) {
      L->warn("matrix expects 16 values, got " + std::to_string(vals.size()) + "; ignored");
      return false;
    }
    *out = Affine{{vals[0], vals[1], vals[2], vals[4], vals[5], vals[6], vals[8], vals[9], vals[10]},
                  {vals[3], vals[7], vals[11]}};  // Transform::from_matrix4 (transform.rs:60-67)
    return true;
  }
  if (el.tag == "lookat") return false;  // only the sensor reads lookat
  L->warn("unsupported transform op <" + el.tag + ">, ignored");
  return false;
}

Affine parse_transform(const Element& el, tpt_loaded_scene* L) {  // mitsuba.rs:483-491
  Affine acc = identity();
  for (const Element& c : el.children) {
    Affine op;
    if (transform_op(c, L, &op)) acc = compose(acc, op);
  }
  return acc;
}

void parse_shape(const Element& el, const std::string& base, tpt_loaded_scene* L) {  // mitsuba.rs:352-411
  int kind;
  double p[4];
  if (const Element* em = el.child_tag("emitter")) {  // parse_emitter (mitsuba.rs:549-556)
    V3 c;
    if (!el_color(*em, "radiance", &c)) c = V3{1, 1, 1};
    if (em->typ() != "area") L->warn("emitter type '" + em->typ() + "' treated as area light");
    kind = TPT_MAT_DIFFUSE_LIGHT;
    p[0] = c.x; p[1] = c.y; p[2] = c.z; p[3] = 0.0;
  } else if (const Element* b = el.child_tag("bsdf")) {
    parse_bsdf(*b, L, &kind, p);
  } else {
    L->warn("shape type '" + el.typ() + "' without bsdf or emitter; defaulting to diffuse");
    kind = TPT_MAT_LAMBERT;
    p[0] = p[1] = p[2] = 0.5;
    p[3] = 0.0;
  }
  const std::string t = el.typ();
  std::vector<double> tris;
  if (t == "sphere") {
    V3 c;
    if (!el_point(el, "center", &c)) c = V3{0, 0, 0};
    double r;
    if (!el_float(el, "radius", &r)) r = 1.0;
    int m = L->add_material(kind, p[0], p[1], p[2], p[3]);
    L->sph_c.insert(L->sph_c.end(), {c.x, c.y, c.z});
    L->sph_r.push_back(r);
    L->sph_mat.push_back(m);
    return;
  } else if (t == "rectangle") {
    tris = rectangle_tris();
  } else if (t == "cube") {
    tris = cube_tris();
  } else if (t == "disk") {
    tris = disk_tris();
  } else if (t == "obj") {
    const std::string* f = el_string(el, "filename");
    if (!f) {
      L->warn("obj shape without filename; skipped");
      return;
    }
    std::string path = resolve(base, *f), err;
    if (!load_obj(path, &tris, &err)) {
      L->warn("failed to load obj '" + path + "': " + err + "; skipped");
      return;
    }
  } else {
    L->warn("unsupported shape type '" + t + "', skipped");
    return;
  }
  const Element* te = el.child_tag("transform");
  Affine x = te ? parse_transform(*te, L) : identity();
  int m = L->add_material(kind, p[0], p[1], p[2], p[3]);
  const int mesh = (int)L->mesh_begin.size() - 1;
  L->tri_v.insert(L->tri_v.end(), tris.begin(), tris.end());
  L->tri_mat.insert(L->tri_mat.end(), tris.size() / 9, m);
  L->mesh_begin.push_back((int64_t)(L->tri_v.size() / 9));
  L->inst_mesh.push_back(mesh);
  L->inst_xform.insert(L->inst_xform.end(), x.a, x.a + 9);
  L->inst_xform.insert(L->inst_xform.end(), x.t, x.t + 3);
  L->inst_override.push_back(-1);
}

// scene-level emitter (mitsuba.rs:226-253); false: no environment from this element
bool parse_scene_emitter(const Element& el, const std::string& base, tpt_loaded_scene* L, int* w, int* h,
                         std::vector<double>* rgb) {
  if (el.child_tag("transform")) L->warn("envmap to_world rotation is unsupported; ignored");
  double scale;
  if (!el_float(el, "scale", &scale)) scale = 1.0;
  const std::string t = el.typ();
  if (t == "envmap") {
    const std::string* f = el_string(el, "filename");
    if (!f) return false;
    std::string path = resolve(base, *f), err;
    std::string low = path;
    for (char& c : low) c = (char)std::tolower((unsigned char)c);
    bool ok = false;
    if (low.size() >= 4 && low.compare(low.size() - 4, 4, ".exr") == 0) err = "OpenEXR environment maps are not read by this build";
    else ok = load_hdr(path, w, h, rgb, &err);
    if (!ok) {
      L->warn("failed to load envmap '" + path + "': " + err + "; ignored");
      return false;
    }
  } else if (t == "constant") {
    V3 c;
    if (!el_color(el, "radiance", &c)) c = V3{1, 1, 1};
    *w = *h = 1;
    *rgb = {c.x, c.y, c.z};
  } else {
    L->warn("unsupported scene emitter type '" + t + "', ignored");
    return false;
  }
  if (scale != 1.0)  // EnvMap::scaled (env.rs:52-58)
    for (double& v : *rgb) v = v * scale;
  return true;
}

}  // namespace

extern "C" {

int tpt_scene_load_xml(const char* path, tpt_render_config* cfg, tpt_loaded_scene** out) {
  if (!path || !cfg || !out) return set_last_error(TPT_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  try {
    std::string xml, err;
    if (!read_file(path, &xml, &err)) return set_last_error(TPT_ERR_INVALID_ARGUMENT, err);
    Element root;
    XmlReader rd(xml);
    if (!rd.parse(&root)) return set_last_error(TPT_ERR_INVALID_ARGUMENT, "malformed XML: " + rd.err);
    if (root.tag != "scene") return set_last_error(TPT_ERR_INVALID_ARGUMENT, "root element is not <scene>");
    std::unique_ptr<tpt_loaded_scene> L(new tpt_loaded_scene());
    // first pass: film / sampler / integrator into the configuration (mitsuba.rs:143-180)
    for (const Element& c : root.children) {
      uint64_t v;
      if (c.tag == "film") {
        if (el_int(c, "width", &v)) cfg->width = (int32_t)std::min<uint64_t>(std::max<uint64_t>(v, 1), INT32_MAX);
        if (el_int(c, "height", &v)) cfg->height = (int32_t)std::min<uint64_t>(std::max<uint64_t>(v, 1), INT32_MAX);
      } else if (c.tag == "sampler") {
        if (el_int(c, "sample_count", &v)) cfg->spp = (int32_t)std::min<uint64_t>(std::max<uint64_t>(v, 1), INT32_MAX);
      } else if (c.tag == "integrator") {
        if (el_int(c, "max_depth", &v)) cfg->max_bounces = (int32_t)std::min<uint64_t>(std::max<uint64_t>(v, 1), INT32_MAX);
        if (el_int(c, "rr_depth", &v)) cfg->rr_start = (int32_t)std::min<uint64_t>(v, INT32_MAX);
      }
    }
    if (cfg->width <= 0 || cfg->height <= 0)  // the camera's aspect ratio comes from the film size
      return set_last_error(TPT_ERR_INVALID_ARGUMENT,
                            "film width and height must be positive: set them in cfg or in <film>");
    const std::string base = dir_of(path);
    const double aspect = (double)cfg->width / (double)cfg->height;
    bool have_cam = false, have_env = false;
    tpt_camera_desc cam{};
    int env_w = 1, env_h = 1;
    std::vector<double> env{0.0, 0.0, 0.0};
    for (const Element& c : root.children) {
      if (c.tag == "sensor") {
        if (c.typ() != "perspective") {
          L->warn("unsupported sensor type '" + c.typ() + "', ignored");
          continue;
        }
        // parse_sensor (mitsuba.rs:307-337)
        double fov;
        if (!el_float(c, "fov", &fov)) fov = 40.0;
        const std::string* ax = el_string(c, "fov_axis");
        double vfov = fov;
        if (!(ax && *ax == "y")) {  // horizontal field of view -> vertical
          double hr = fov * (kPi / 180.0);
          vfov = (2.0 * std::atan(std::tan(hr * 0.5) / aspect)) * (180.0 / kPi);
        }
        V3 eye{0, 0, 4}, target{0, 0, 0}, up{0, 1, 0};
        const Element* tr = c.child_tag("transform");
        const Element* la = tr ? tr->child_tag("lookat") : nullptr;
        if (la) {  // parse_lookat (mitsuba.rs:339-349)
          const std::string *o = la->attr("origin"), *t = la->attr("target"), *u = la->attr("up");
          V3 eo, to;
          if (o && t && parse_vec3(*o, &eo) && parse_vec3(*t, &to)) {
            eye = eo;
            target = to;
            V3 uo;
            up = (u && parse_vec3(*u, &uo)) ? uo : V3{0, 1, 0};
          }
        }
        double ar;
        if (!el_float(c, "aperture_radius", &ar)) ar = 0.0;
        double focus;
        if (!el_float(c, "focus_distance", &focus)) {
          double dx = eye.x - target.x, dy = eye.y - target.y, dz = eye.z - target.z;
          focus = std::sqrt(dx * dx + dy * dy + dz * dz);
        }
        cam = tpt_camera_desc{{eye.x, eye.y, eye.z}, {target.x, target.y, target.z}, {up.x, up.y, up.z}, vfov, aspect,
                              focus, 2.0 * ar};
        have_cam = true;
      } else if (c.tag == "shape") {
        parse_shape(c, base, L.get());
      } else if (c.tag == "emitter") {
        int w = 0, h = 0;
        std::vector<double> rgb;
        if (parse_scene_emitter(c, base, L.get(), &w, &h, &rgb)) {
          env_w = w;
          env_h = h;
          env = std::move(rgb);
          have_env = true;
        }
      } else if (c.tag == "integrator" || c.tag == "sampler" || c.tag == "film" || c.tag == "default" ||
                 c.tag == "rfilter") {
      } else {
        L->warn("unsupported element <" + c.tag + ">, skipped");
      }
    }
    if (!have_cam) {
      L->warn("no perspective sensor found; using a default camera");
      cam = tpt_camera_desc{{0, 0, 4}, {0, 0, 0}, {0, 1, 0}, 40.0, aspect, 4.0, 0.0};  // Camera::look_at
    }
    (void)have_env;  // no environment emitter: a black 1x1 map (mitsuba.rs:216-218)
    L->env_rgb = std::move(env);
    L->desc.camera = cam;
    L->desc.env_kind = TPT_ENV_IMAGE;
    L->desc.env_width = env_w;
    L->desc.env_height = env_h;
    L->finish();
    *out = L.release();
    return TPT_OK;
  } catch (const std::bad_alloc&) {
    return set_last_error(TPT_ERR_OUT_OF_MEMORY, "host allocation failed while loading the scene");
  } catch (const std::exception& e) {
    return set_last_error(TPT_ERR_INTERNAL, e.what());
  }
}

int tpt_loaded_scene_desc(const tpt_loaded_scene* scene, const tpt_scene_desc** desc) {
  if (!scene || !desc) return set_last_error(TPT_ERR_INVALID_ARGUMENT, "null argument");
  *desc = &scene->desc;
  return TPT_OK;
}

const char* tpt_loaded_scene_warnings(const tpt_loaded_scene* scene) { return scene ? scene->warnings.c_str() : ""; }

int tpt_loaded_scene_destroy(tpt_loaded_scene* scene) {
  delete scene;
  return TPT_OK;
}

int tpt_obj_load(const char* path, int64_t* n_tris, double* tri_vertices) {
  if (!path || !n_tris) return set_last_error(TPT_ERR_INVALID_ARGUMENT, "null argument");
  try {
    std::vector<double> tris;
    std::string err;
    if (!load_obj(path, &tris, &err)) return set_last_error(TPT_ERR_INVALID_ARGUMENT, err);
    const int64_t n = (int64_t)(tris.size() / 9), cap = *n_tris;
    *n_tris = n;
    if (tri_vertices && n > cap)
      return set_last_error(TPT_ERR_INVALID_ARGUMENT, "the buffer holds " + std::to_string(cap) +
                                                          " triangles, the file " + std::to_string(n));
    if (tri_vertices && n > 0) std::memcpy(tri_vertices, tris.data(), sizeof(double) * tris.size());
    return TPT_OK;
  } catch (const std::bad_alloc&) {
    return set_last_error(TPT_ERR_OUT_OF_MEMORY, "host allocation failed while reading the OBJ file");
  }
}

}  // extern "C"
