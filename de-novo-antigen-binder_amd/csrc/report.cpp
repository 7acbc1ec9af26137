// Candidate report writers (src/report.rs:41-113): FASTA records, a JSON
// array and CSV rows of the designed candidates' summaries, formatted as the
// reference formats them. Text out; the caller writes the file.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "dnab_internal.h"

namespace dnab {
namespace {

// Rust's {:.N} of an f32: the exact value rounded half to even to N places,
// as printf's %.Nf does; NaN and infinities as Rust spells them.
void put_fixed(std::string& out, float v, int places) {
  char buf[64];
  if (std::isnan(v)) {
    out += "NaN";
    return;
  }
  if (std::isinf(v)) {
    out += v > 0 ? "inf" : "-inf";
    return;
  }
  std::snprintf(buf, sizeof(buf), "%.*f", places, static_cast<double>(v));
  out += buf;
}

void put_uint(std::string& out, unsigned long long v) { out += std::to_string(v); }

// json_escape (src/report.rs:22-34): quote, backslash and newline only
void put_json_string(std::string& out, const char* s) {
  out += '"';
  for (const char* c = s; *c; ++c) {
    if (*c == '"') out += "\\\"";
    else if (*c == '\\') out += "\\\\";
    else if (*c == '\n') out += "\\n";
    else out += *c;
  }
  out += '"';
}

// csv_field (src/report.rs:36-39): quoted, embedded quotes doubled
void put_csv_field(std::string& out, const char* s) {
  out += '"';
  for (const char* c = s; *c; ++c) {
    if (*c == '"') out += '"';
    out += *c;
  }
  out += '"';
}

void fasta(std::string& out, const dnab_candidate_report& r) {
  const std::string rank = std::to_string(r.rank);
  if (r.light_sequence) {
    out += ">candidate_" + rank + "_VH E=";
    put_fixed(out, r.energy, 4);
    out += '\n';
    out += r.sequence;
    out += "\n>candidate_" + rank + "_VL E=";
    put_fixed(out, r.energy, 4);
    out += '\n';
    out += r.light_sequence;
    out += '\n';
  } else {
    out += ">candidate_" + rank + " E=";
    put_fixed(out, r.energy, 4);
    out += '\n';
    out += r.sequence;
    out += '\n';
  }
}

void json(std::string& out, const dnab_candidate_report& r, bool last) {
  out += "  {\n    \"rank\": ";
  put_uint(out, r.rank);
  out += ",\n    \"sequence\": ";
  put_json_string(out, r.sequence);
  out += ",\n    \"light_sequence\": ";
  if (r.light_sequence) put_json_string(out, r.light_sequence);
  else out += "null";
  out += ",\n    \"energy\": ";
  put_fixed(out, r.energy, 6);
  out += ",\n    \"entropy_penalty\": ";
  put_fixed(out, r.entropy_penalty, 6);
  out += ",\n    \"dg_corrected\": ";
  put_fixed(out, r.dg_corrected, 6);
  out += ",\n    \"net_charge\": " + std::to_string(r.net_charge);
  out += ",\n    \"aggregation_risk\": ";
  out += r.aggregation_risk ? "true" : "false";
  out += ",\n    \"n_interface\": ";
  put_uint(out, r.n_interface);
  out += ",\n    \"clashes\": ";
  out += r.clashes < 0 ? std::string("null") : std::to_string(r.clashes);
  out += ",\n    \"rama_outliers\": ";
  out += r.rama_outliers < 0 ? std::string("null") : std::to_string(r.rama_outliers);
  out += last ? "\n  }\n" : "\n  },\n";
}

void csv(std::string& out, const dnab_candidate_report& r) {
  put_uint(out, r.rank);
  out += ',';
  put_csv_field(out, r.sequence);
  out += ',';
  put_csv_field(out, r.light_sequence ? r.light_sequence : "");
  for (float v : {r.energy, r.entropy_penalty, r.dg_corrected}) {
    out += ',';
    put_fixed(out, v, 6);
  }
  out += ',' + std::to_string(r.net_charge);
  out += r.aggregation_risk ? ",true," : ",false,";
  put_uint(out, r.n_interface);
  out += ',';
  if (r.clashes >= 0) out += std::to_string(r.clashes);
  out += ',';
  if (r.rama_outliers >= 0) out += std::to_string(r.rama_outliers);
  out += '\n';
}

}  // namespace
}  // namespace dnab

using namespace dnab;

extern "C" int dnab_write_report(const dnab_candidate_report* reports, size_t n, int format,
                                 char* out, size_t capacity, size_t* n_chars) {
  if (n && !reports) return DNAB_ERR_INVALID;
  if (format != DNAB_REPORT_FASTA && format != DNAB_REPORT_JSON && format != DNAB_REPORT_CSV)
    return DNAB_ERR_INVALID;
  for (size_t i = 0; i < n; ++i)
    if (!reports[i].sequence) return DNAB_ERR_INVALID;
  std::string s;
  if (format == DNAB_REPORT_JSON) s = "[\n";
  if (format == DNAB_REPORT_CSV)
    s = "rank,sequence,light_sequence,energy,entropy_penalty,dg_corrected,net_charge,"
        "aggregation_risk,n_interface,clashes,rama_outliers\n";
  for (size_t i = 0; i < n; ++i) {
    if (format == DNAB_REPORT_FASTA) fasta(s, reports[i]);
    else if (format == DNAB_REPORT_JSON) json(s, reports[i], i + 1 == n);
    else csv(s, reports[i]);
  }
  if (format == DNAB_REPORT_JSON) s += "]\n";
  if (n_chars) *n_chars = s.size();
  if (!out) return DNAB_OK;
  if (s.size() + 1 > capacity) return DNAB_ERR_CAPACITY;
  std::memcpy(out, s.c_str(), s.size() + 1);
  return DNAB_OK;
}
