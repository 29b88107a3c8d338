// swm_namelist.cpp -- the run settings file of the reference driver,
// data/namelist, as model_control.F:98-139 (input) reads it: three F7.2
// reals, five L1 logicals and three list-directed integers, one per line,
// each followed by a comment.
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>

#include "../../include/swm_mi355x.h"
#include "swm_internal.h"

namespace {

// Fortran F7.2 input of the first 7 characters: blanks are ignored; without
// a decimal point the last 2 digits are the fraction
bool read_f72(const std::string& line, double* out) {
  std::string f = line.substr(0, 7), t;
  for (char ch : f)
    if (ch != ' ') t += ch;
  if (t.empty()) { *out = 0.0; return true; }
  char* end = nullptr;
  const double v = std::strtod(t.c_str(), &end);
  if (!end || *end != '\0') return false;
  *out = t.find_first_of(".eEdD") == std::string::npos ? v / 100.0 : v;
  return true;
}

// L1 input: T or F in the first column
bool read_l1(const std::string& line, int* out) {
  if (line.empty()) return false;
  const char ch = (char)std::toupper((unsigned char)line[0]);
  if (ch != 'T' && ch != 'F') return false;
  *out = ch == 'T';
  return true;
}

// list-directed integer: the first token
bool read_int(const std::string& line, int* out) {
  const char* s = line.c_str();
  while (*s == ' ' || *s == '\t') ++s;
  char* end = nullptr;
  const long v = std::strtol(s, &end, 10);
  if (end == s || (*end != '\0' && *end != ' ' && *end != '\t' && *end != ',' && *end != '!' && *end != '\r')) return false;
  *out = (int)v;
  return true;
}

}  // namespace

extern "C" swm_status swm_read_namelist(const char* path, swm_namelist* nl) {
  if (!path || !nl) { swm::set_error("swm_read_namelist: bad argument"); return SWM_ERR_ARG; }
  std::ifstream in(path);
  if (!in) { swm::set_error(std::string("swm_read_namelist: cannot open ") + path); return SWM_ERR_ARG; }
  std::string ln[11];
  for (int k = 0; k < 11; ++k)
    if (!std::getline(in, ln[k])) { swm::set_error("swm_read_namelist: fewer than 11 lines in " + std::string(path)); return SWM_ERR_ARG; }
  swm_namelist v{};
  const char* bad = nullptr;
  if (!read_f72(ln[0], &v.tau_end_day)) bad = "tau_end_day";
  else if (!read_f72(ln[1], &v.tau_end_hour)) bad = "tau_end_hour";
  else if (!read_f72(ln[2], &v.dt)) bad = "dt";
  else if (!read_l1(ln[3], &v.l_restart)) bad = "l_restart";
  else if (!read_l1(ln[4], &v.l_upwind)) bad = "l_upwind";
  else if (!read_l1(ln[5], &v.l_centered)) bad = "l_centered";
  else if (!read_l1(ln[6], &v.l_fct)) bad = "l_fct";
  else if (!read_l1(ln[7], &v.l_remap)) bad = "l_remap";
  else if (!read_int(ln[8], &v.n_restart)) bad = "n_restart";
  else if (!read_int(ln[9], &v.n_output)) bad = "n_output";
  else if (!read_int(ln[10], &v.n_diagnostics)) bad = "n_diagnostics";
  if (bad) { swm::set_error(std::string("swm_read_namelist: cannot read ") + bad + " in " + path); return SWM_ERR_ARG; }
  *nl = v;
  return SWM_OK;
}
