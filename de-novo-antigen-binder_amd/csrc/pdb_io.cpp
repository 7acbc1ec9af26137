// PDB text I/O around the path: the antigen comes in as the Calpha atoms of a
// PDB file (parse_pdb, src/pdb.rs:19-57) and results go out as fixed-column
// ATOM records (write_pdb, write_pdb_allatom, write_pdb_multi,
// src/pdb.rs:94-191). Text in, text out: the caller does the file I/O.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "dnab_internal.h"
#include "ff_tables.h"

namespace dnab {
namespace {

const char* const kThreeLetter[kNumAminoAcids] = {"ALA", "ARG", "ASN", "ASP", "CYS", "GLN", "GLU",
                                                  "GLY", "HIS", "ILE", "LEU", "LYS", "MET", "PHE",
                                                  "PRO", "SER", "THR", "TRP", "TYR", "VAL"};

// AminoAcid::from_three_letter (src/atom.rs:127-151): HID/HIE/HIP are His,
// anything unknown is Gly.
uint8_t from_three_letter(const std::string& s) {
  for (int a = 0; a < kNumAminoAcids; ++a)
    if (s == kThreeLetter[a]) return static_cast<uint8_t>(a);
  if (s == "HID" || s == "HIE" || s == "HIP") return kHis;
  return kGly;
}

std::string trim(const std::string& s) {
  size_t b = 0, e = s.size();
  while (b < e && std::isspace(static_cast<unsigned char>(s[b]))) ++b;
  while (e > b && std::isspace(static_cast<unsigned char>(s[e - 1]))) --e;
  return s.substr(b, e - b);
}

// line.get(a..b): the bytes a..b, or nothing when the line is shorter
bool field(const std::string& line, size_t a, size_t b, std::string& out) {
  if (line.size() < b) return false;
  out = line.substr(a, b - a);
  return true;
}

// str::parse::<f32>: a decimal number (sign, digits, point, exponent) or
// inf / infinity / nan, correctly rounded; anything else fails.
bool parse_f32(const std::string& s, float& v) {
  if (s.empty()) return false;
  for (char c : s)
    if (c == 'x' || c == 'X' || c == '(' || std::isspace(static_cast<unsigned char>(c)))
      return false;  // no hexadecimal forms, no nan payloads, no inner spaces
  errno = 0;
  char* end = nullptr;
  const float f = std::strtof(s.c_str(), &end);
  if (end != s.c_str() + s.size()) return false;
  v = f;
  return true;
}

// {:8.3} of an f32, as Rust prints it (NaN, inf, -inf for the specials)
void put_coord(std::string& out, float v) {
  char buf[64];
  if (std::isnan(v)) std::snprintf(buf, sizeof(buf), "%8s", "NaN");
  else if (std::isinf(v)) std::snprintf(buf, sizeof(buf), "%8s", v > 0 ? "inf" : "-inf");
  else std::snprintf(buf, sizeof(buf), "%8.3f", static_cast<double>(v));
  out += buf;
}

void put_atom(std::string& out, size_t serial, const char* name, uint8_t aa, char chain,
              size_t res_seq, float x, float y, float z) {
  // "ATOM  {:5} {:<4} {:3} {:1}{:4}    {:8.3}{:8.3}{:8.3}  1.00  0.00           C\n"
  char buf[96];
  std::snprintf(buf, sizeof(buf), "ATOM  %5zu %-4s %-3s %c%4zu    ", serial < 99999 ? serial : 99999,
                name, kThreeLetter[aa], chain, res_seq < 9999 ? res_seq : 9999);
  out += buf;
  put_coord(out, x);
  put_coord(out, y);
  put_coord(out, z);
  out += "  1.00  0.00           C\n";
}

int emit(const std::string& s, char* out, size_t capacity, size_t* n_chars) {
  if (n_chars) *n_chars = s.size();
  if (!out) return DNAB_OK;  // size query
  if (s.size() + 1 > capacity) return DNAB_ERR_CAPACITY;
  std::memcpy(out, s.c_str(), s.size() + 1);
  return DNAB_OK;
}

bool chain_ok(const dnab_pdb_chain& c) {
  if (c.n_res && (!c.res_start || !c.amino_acid)) return false;
  for (size_t r = 0; r < c.n_res; ++r)
    if (c.amino_acid[r] >= kNumAminoAcids || c.res_start[r + 1] < c.res_start[r]) return false;
  const size_t na = c.n_res ? c.res_start[c.n_res] : 0;
  return na == 0 || (c.x && c.y && c.z);
}

// the ATOM records of one all-atom chain (src/pdb.rs:121-155)
void chain_records(std::string& out, const dnab_pdb_chain& c, size_t& serial) {
  for (size_t r = 0; r < c.n_res; ++r) {
    const uint8_t aa = c.amino_acid[r];
    const ResTopology& topo = kResidueTopology[aa];
    for (uint32_t k = 0, i = c.res_start[r]; i < c.res_start[r + 1]; ++k, ++i) {
      if (k >= topo.n_atoms) break;
      ++serial;
      put_atom(out, serial, topo.atoms[k].name, aa, c.chain_id, r + 1, c.x[i], c.y[i], c.z[i]);
    }
  }
}

}  // namespace
}  // namespace dnab

using namespace dnab;

extern "C" {

int dnab_parse_pdb(const char* text, size_t len, float* x, float* y, float* z,
                   uint8_t* amino_acid, size_t capacity, size_t* n_out) {
  if (!text && len) return DNAB_ERR_INVALID;
  size_t n = 0;
  size_t pos = 0;
  while (pos < len) {
    size_t e = pos;
    while (e < len && text[e] != '\n') ++e;
    std::string line(text + pos, e - pos);
    if (!line.empty() && line.back() == '\r') line.pop_back();
    pos = e + 1;
    if (line.compare(0, 4, "ATOM") != 0 && line.compare(0, 6, "HETATM") != 0) continue;
    std::string f;
    const std::string name = field(line, 12, 16, f) ? trim(f) : std::string();
    if (name != "CA") continue;
    const std::string res = field(line, 17, 20, f) ? trim(f) : std::string();
    float c[3] = {0.0f, 0.0f, 0.0f};
    const size_t cols[3][2] = {{30, 38}, {38, 46}, {46, 54}};
    for (int k = 0; k < 3; ++k) {
      float v;
      if (field(line, cols[k][0], cols[k][1], f) && parse_f32(trim(f), v)) c[k] = v;
    }
    if (n < capacity) {
      if (!x || !y || !z || !amino_acid) return DNAB_ERR_INVALID;
      x[n] = c[0];
      y[n] = c[1];
      z[n] = c[2];
      amino_acid[n] = from_three_letter(res);
    }
    ++n;
  }
  if (n_out) *n_out = n;
  if (n == 0) return DNAB_ERR_INVALID;  // BinderError::EmptyInput
  return n > capacity ? DNAB_ERR_CAPACITY : DNAB_OK;
}

int dnab_write_pdb_ca(const float* x, const float* y, const float* z, const uint8_t* amino_acid,
                      size_t n, char chain_id, char* out, size_t capacity, size_t* n_chars) {
  if (n && (!x || !y || !z || !amino_acid)) return DNAB_ERR_INVALID;
  std::string s;
  s.reserve(n * 81 + 4);
  for (size_t i = 0; i < n; ++i) {
    if (amino_acid[i] >= kNumAminoAcids) return DNAB_ERR_INVALID;
    put_atom(s, i + 1, " CA ", amino_acid[i], chain_id, i + 1, x[i], y[i], z[i]);
  }
  s += "END\n";
  return emit(s, out, capacity, n_chars);
}

int dnab_write_pdb_chains(const dnab_pdb_chain* chains, size_t n_chains, int ter, char* out,
                          size_t capacity, size_t* n_chars) {
  if (n_chains && !chains) return DNAB_ERR_INVALID;
  std::string s;
  size_t serial = 0;
  for (size_t c = 0; c < n_chains; ++c) {
    if (!chain_ok(chains[c])) return DNAB_ERR_INVALID;
    chain_records(s, chains[c], serial);
    if (ter) s += "TER\n";
  }
  s += "END\n";
  return emit(s, out, capacity, n_chars);
}

}  // extern "C"
