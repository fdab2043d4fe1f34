// Host parsers for the large ASCII lists of an OpenFOAM polyMesh (points,
// faces, owner / neighbour), the inner loops of io/mesh_io.py:154-369 in
// one pass over the bytes instead of a regular expression per entry.
//
// Each entry point takes the text that follows the outer list's opening
// parenthesis (reading stops at its closing one, or at the end of the text)
// and the count the file states, and reproduces the reference's reading of
// the list (which first finds the matching parenthesis, :617-627):
//   points  every "( ... )" group holds the coordinates of one point
//           (:154-171; here exactly three numbers per group);
//   faces   every "<k>( ... )" group holds one face's vertex labels, the
//           count k itself unchecked, as the reference's regex leaves it
//           (:255-274);
//   labels  whitespace-separated integers (:352-366).
// Numbers are read with strtod / strtoll (correctly rounded, like Python's
// float()).  A count different from the stated one is an error, as in the
// reference.

#include <cerrno>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>

#include "pfm_internal.h"

namespace {

inline bool is_space(char c) { return c == ' ' || c == '\n' || c == '\r' || c == '\t' || c == '\f' || c == '\v'; }

// Null-terminated copy so that strtod cannot read past the slice.
struct Text {
  std::string s;
  Text(const char *p, int64_t n) : s(p, (size_t)n) {}
  const char *begin() const { return s.c_str(); }
  const char *end() const { return s.c_str() + s.size(); }
};

}  // namespace

extern "C" {

int pfm_parse_ascii_points(const char *text, int64_t len, int64_t n, double *out,
                           int64_t *n_read) {
  PFM_REQUIRE(len >= 0 && n >= 0 && n_read, "bad arguments");
  PFM_REQUIRE(len == 0 || text, "null text");
  PFM_REQUIRE(n == 0 || out, "null output");
  Text t(text, len);
  const char *p = t.begin(), *e = t.end();
  int64_t k = 0;
  while (true) {
    while (p < e && *p != '(' && *p != ')') ++p;
    if (p >= e || *p == ')') break;  // end of the outer list
    ++p;
    double v[3];
    int m = 0;
    while (true) {
      while (p < e && is_space(*p)) ++p;
      PFM_REQUIRE(p < e, "unterminated point %lld", (long long)k);
      if (*p == ')') break;
      char *q = nullptr;
      errno = 0;
      const double x = std::strtod(p, &q);
      PFM_REQUIRE(q != p, "point %lld: not a number near '%.16s'", (long long)k, p);
      PFM_REQUIRE(m < 3, "point %lld has more than 3 coordinates", (long long)k);
      v[m++] = x;
      p = q;
    }
    ++p;  // ')'
    PFM_REQUIRE(m == 3, "point %lld has %d coordinates", (long long)k, m);
    if (k < n) {
      out[3 * k] = v[0];
      out[3 * k + 1] = v[1];
      out[3 * k + 2] = v[2];
    }
    ++k;
  }
  *n_read = k;
  PFM_REQUIRE(k == n, "Expected %lld points, got %lld", (long long)n, (long long)k);
  return PFM_OK;
}

int pfm_parse_ascii_faces(const char *text, int64_t len, int64_t n, int64_t verts_cap,
                          int32_t *face_ptr, int32_t *verts, int64_t *n_read,
                          int64_t *n_verts) {
  PFM_REQUIRE(len >= 0 && n >= 0 && verts_cap >= 0 && n_read && n_verts, "bad arguments");
  PFM_REQUIRE(len == 0 || text, "null text");
  PFM_REQUIRE(face_ptr, "null face_ptr");
  Text t(text, len);
  const char *p = t.begin(), *e = t.end();
  int64_t k = 0, nv = 0;
  face_ptr[0] = 0;
  while (true) {
    // a face: digits immediately followed by '('; ')' ends the outer list
    while (p < e && !(*p >= '0' && *p <= '9') && *p != ')') ++p;
    if (p >= e || *p == ')') break;
    while (p < e && *p >= '0' && *p <= '9') ++p;
    if (p >= e || *p != '(') continue;  // a number not followed by '(' is not a face
    ++p;
    while (true) {
      while (p < e && is_space(*p)) ++p;
      PFM_REQUIRE(p < e, "unterminated face %lld", (long long)k);
      if (*p == ')') break;
      char *q = nullptr;
      const long long v = std::strtoll(p, &q, 10);
      PFM_REQUIRE(q != p, "face %lld: not a label near '%.16s'", (long long)k, p);
      PFM_REQUIRE(v >= INT32_MIN && v <= INT32_MAX, "face %lld: label out of range",
                  (long long)k);
      if (nv < verts_cap && verts) verts[nv] = (int32_t)v;
      ++nv;
      p = q;
    }
    ++p;
    ++k;
    PFM_REQUIRE(nv <= INT32_MAX, "too many face vertices");
    if (k <= n) face_ptr[k] = (int32_t)nv;
  }
  *n_read = k;
  *n_verts = nv;
  PFM_REQUIRE(k == n, "Expected %lld faces, got %lld", (long long)n, (long long)k);
  PFM_REQUIRE(nv <= verts_cap, "vertex buffer holds %lld labels, the faces have %lld",
              (long long)verts_cap, (long long)nv);
  return PFM_OK;
}

int pfm_parse_ascii_labels(const char *text, int64_t len, int64_t n, int64_t *out,
                           int64_t *n_read) {
  PFM_REQUIRE(len >= 0 && n >= 0 && n_read, "bad arguments");
  PFM_REQUIRE(len == 0 || text, "null text");
  PFM_REQUIRE(n == 0 || out, "null output");
  Text t(text, len);
  const char *p = t.begin(), *e = t.end();
  int64_t k = 0;
  while (true) {
    while (p < e && is_space(*p)) ++p;
    if (p >= e || *p == ')') break;  // end of the outer list
    char *q = nullptr;
    const long long v = std::strtoll(p, &q, 10);
    PFM_REQUIRE(q != p && (q == e || is_space(*q) || *q == ')'),
                "label %lld: not an integer near '%.16s'", (long long)k, p);
    if (k < n) out[k] = v;
    ++k;
    p = q;
  }
  *n_read = k;
  PFM_REQUIRE(k == n, "Expected %lld labels, got %lld", (long long)n, (long long)k);
  return PFM_OK;
}

}  // extern "C"
