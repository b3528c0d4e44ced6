// read.input ingestion: the keyword blocks of READIN0 (src/mod_init.f90:401-710)
// that set the grid, truncation and physics, mapped onto mlg_params.
//
// Line handling follows READCOM1 (src/mod_init.f90:165-372): a line whose
// first character is '#' is a comment, '!' lines (shell commands in the
// reference) are skipped, '%END' ends the file, commas and tabs separate
// fields like blanks, and "..." / '...' / (...) group one field.  Numbers
// are read like ATOF's F72.0 (src/mod_misc.f90:1622-1657): D or E exponents,
// an unreadable field gives 0; integer settings take the value truncated
// toward zero, as the REAL-to-INTEGER assignments do.
//
// Quirks kept: ZLEN <= 1 is a wavenumber (ZLEN = 2 pi / value); VISCOSITY
// NU or NUP >= 1 is a Reynolds number (value = 1/value); BOUSSINESQ's fifth
// field > 0 selects AB2-CN.  Blocks of the initial-condition, output,
// diagnostics and post-processing settings are recognised and skipped.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "mlegs_hip.h"

namespace {

std::vector<std::string> fields(std::string c) {
  for (char &ch : c)
    if (ch == ',' || ch == '\t' || ch == '\r') ch = ' ';
  std::vector<std::string> out;
  size_t i = 0;
  while (i < c.size()) {
    while (i < c.size() && c[i] == ' ') i++;
    if (i >= c.size()) break;
    char q = c[i];
    char close = q == '"' ? '"' : q == '\'' ? '\'' : q == '(' ? ')' : 0;
    if (close) {
      size_t j = c.find(close, i + 1);
      if (j == std::string::npos) j = c.size();
      out.push_back(c.substr(i + 1, j - i - 1));
      i = j + 1;
    } else {
      size_t j = c.find(' ', i);
      if (j == std::string::npos) j = c.size();
      out.push_back(c.substr(i, j - i));
      i = j;
    }
  }
  return out;
}

double atof_f(const std::string &s) {
  std::string t;
  for (char ch : s) t += (ch == 'd' || ch == 'D') ? 'E' : ch;
  if (t.empty()) return 0.0;
  char *end = nullptr;
  double v = std::strtod(t.c_str(), &end);
  if (end == t.c_str()) return 0.0;
  return v;
}

int atoi_f(const std::string &s) { return (int)std::trunc(atof_f(s)); }

struct Reader {
  std::ifstream in;
  bool done = false;
  // next non-comment line split into fields ({} at end of file)
  bool next(std::vector<std::string> &f) {
    std::string line;
    while (!done && std::getline(in, line)) {
      if (line.size() > 200) line.resize(200);
      if (line.rfind("%END", 0) == 0) {
        done = true;
        break;
      }
      if (!line.empty() && (line[0] == '#' || line[0] == '!')) continue;
      f = fields(line);
      return true;
    }
    done = true;
    f.clear();
    return false;
  }
};

}  // namespace

extern "C" int mlg_read_input(const char *path, mlg_params *p, double *t_limit) {
  if (!path || !p) return MLG_E_ARG;
  Reader r;
  r.in.open(path);
  if (!r.in) return MLG_E_ARG;
  mlg_params q;
  std::memset(&q, 0, sizeof(q));
  q.minc = 1;
  q.ell = 1.0;
  q.visc_sw = 0;
  double limit = 0.0;
  int nvortex = 1;
  const double pi = std::acos(-1.0);
  std::vector<std::string> f;
  int lines = 0;
  while (r.next(f) && lines++ < 1000) {
    if (f.empty()) continue;
    const std::string &key = f[0];  // keywords are compared as written, like COM(1).EQ.'...'
    std::vector<std::string> v;
    auto row = [&]() {
      if (!r.next(v)) v.clear();
      v.resize(8);
    };
    if (key == "DIMENSION") {
      row();
      q.nr = atoi_f(v[0]);
      q.nth = atoi_f(v[1]);
      q.nx = atoi_f(v[2]);
    } else if (key == "CHOPS") {
      row();
      q.nrchop = atoi_f(v[0]);
      q.ntchop = atoi_f(v[1]);
      q.nxchop = atoi_f(v[2]);
    } else if (key == "ELL") {
      row();
      q.ell = atof_f(v[0]);
    } else if (key == "ZLEN") {
      row();
      double z = atof_f(v[0]);
      q.zlen = z > 1.0 ? z : 2.0 * pi / z;
    } else if (key == "MINC") {
      row();
      q.minc = atoi_f(v[0]);
    } else if (key == "MKLINK") {
      row();
      q.mklink = atoi_f(v[0]);
    } else if (key == "DT") {
      row();
      q.dt = atof_f(v[0]);
      limit = atof_f(v[1]);
    } else if (key == "VISCOSITY") {
      row();
      q.visc_sw = atoi_f(v[0]);
      q.nu = atof_f(v[1]);
      if (q.nu >= 1.0) q.nu = std::pow(q.nu, -1.0);
      q.nup = atof_f(v[2]);
      if (q.nup >= 1.0) q.nup = std::pow(q.nup, -1.0);
      q.visc_p = atoi_f(v[3]);
    } else if (key == "BOUSSINESQ") {
      row();
      q.bv0 = atof_f(v[0]);
      q.omega = atof_f(v[1]);
      q.kappa = atof_f(v[2]);
      q.kappap = atof_f(v[3]);
      q.adams = atof_f(v[4]) > 0 ? 1 : 0;
    } else if (key == "VORTEXPAIR") {
      row();
      nvortex = atoi_f(v[0]);
      for (int i = 0; i < nvortex; i++) row();
    } else if (key == "VORTEXPAIR_CROW" || key == "VORTEXPAIR_SECTION") {
      for (int i = 0; i < nvortex; i++) row();  // one line per vortex
    } else if (key == "OUTPUTS") {
      row();
      int n = atoi_f(v[0]);
      for (int i = 0; i < n; i++) row();
    } else if (key == "MONITOR_MK") {
      row();
      int n = atoi_f(v[0]);
      for (int i = 0; i < n + 1; i++) row();
    } else if (key == "POSTPROCESS") {
      row();
      int n = atoi_f(v[0]);
      for (int i = 0; i < n; i++) row();
    } else if (key == "END") {
      break;
    } else {
      // one-line blocks: VORTEXPAIR_NOISE, TIME_INTEG, NUP_ADJUST, FREESTREAM,
      // STRAIN, REMOVE, SAVEDIR, PSI0, CHI0, B0, PSII, CHII, BI, DIAGNOST,
      // LINEAR, DEBUG; an unknown keyword is skipped, as READIN0 warns and
      // continues
      static const char *one[] = {"VORTEXPAIR_NOISE", "TIME_INTEG", "NUP_ADJUST", "FREESTREAM", "STRAIN", "REMOVE",
                                  "SAVEDIR", "PSI0", "CHI0", "B0", "PSII", "CHII", "BI", "DIAGNOST", "LINEAR", "DEBUG"};
      for (const char *k : one)
        if (key == k) {
          row();
          break;
        }
    }
  }
  *p = q;
  if (t_limit) *t_limit = limit;
  return 0;
}

// The run settings besides the grid and physics: READIN0's SAVEDIR, PSI0,
// CHI0, B0, PSII, CHII, BI, OUTPUTS, DIAGNOST, REMOVE, NUP_ADJUST and
// MONITOR_MK blocks (src/mod_init.f90:551-660), with READIN0's defaults
// (:428-447).  Names are stored blank-padded like CHARACTER(LEN=72).
namespace {
void put_name(char *dst, const std::string &s) {
  std::memset(dst, ' ', MLG_NAME_LEN);
  std::memcpy(dst, s.data(), s.size() < (size_t)MLG_NAME_LEN ? s.size() : (size_t)MLG_NAME_LEN);
}
}  // namespace

extern "C" int mlg_read_run_settings(const char *path, mlg_run_settings *s) {
  if (!path || !s) return MLG_E_ARG;
  Reader r;
  r.in.open(path);
  if (!r.in) return MLG_E_IO;
  std::memset(s, 0, sizeof(*s));
  const char *blank_names[] = {s->savedir, s->psi0, s->chi0, s->b0, s->psii, s->chii, s->bi};
  for (const char *p : blank_names) std::memset((char *)p, ' ', MLG_NAME_LEN);
  s->it1 = 40;
  s->it2 = 40;
  s->rmv_sw = 0;
  s->rmv_int = 10000;
  s->adj_sw = 0;
  int nvortex = 1;
  std::vector<std::string> f, v;
  int lines = 0;
  auto row = [&]() {
    if (!r.next(v)) v.clear();
    v.resize(8);
  };
  while (r.next(f) && lines++ < 1000) {
    if (f.empty()) continue;
    const std::string &key = f[0];
    if (key == "SAVEDIR") {
      row();
      put_name(s->savedir, v[0]);
    } else if (key == "PSI0") {
      row();
      put_name(s->psi0, v[0]);
    } else if (key == "CHI0") {
      row();
      put_name(s->chi0, v[0]);
    } else if (key == "B0") {
      row();
      put_name(s->b0, v[0]);
    } else if (key == "PSII") {
      row();
      put_name(s->psii, v[0]);
    } else if (key == "CHII") {
      row();
      put_name(s->chii, v[0]);
    } else if (key == "BI") {
      row();
      put_name(s->bi, v[0]);
    } else if (key == "OUTPUTS") {
      row();
      const int n = atoi_f(v[0]);
      if (n < 0 || n > MLG_MAX_OUTPUTS) return MLG_E_ARG;
      s->noutputs = n;
      for (int i = 0; i < n; i++) {
        row();
        s->out_t[i] = atof_f(v[0]);
        put_name(s->out_psi[i], v[1]);
        put_name(s->out_chi[i], v[2]);
        put_name(s->out_b[i], v[3]);
      }
    } else if (key == "DIAGNOST") {
      row();
      s->it1 = atoi_f(v[0]);
      s->it2 = atoi_f(v[1]);
    } else if (key == "REMOVE") {
      row();
      s->rmv_sw = atoi_f(v[0]);
      s->rmv_int = atoi_f(v[1]);
    } else if (key == "NUP_ADJUST") {
      row();
      s->adj_sw = atoi_f(v[0]);
      s->adj_int = atoi_f(v[1]);
    } else if (key == "MONITOR_MK") {
      row();
      const int n = atoi_f(v[0]);
      if (n < 0 || n > MLG_MAX_MONITOR) return MLG_E_ARG;
      s->nmonitor = n;
      for (int i = 0; i < n; i++) {
        row();
        s->monitor_mk[i][0] = atoi_f(v[0]);
        s->monitor_mk[i][1] = atoi_f(v[1]);
      }
      row();
      s->monitor_start = atoi_f(v[0]);
      s->monitor_interval = atoi_f(v[1]);
    } else if (key == "VORTEXPAIR") {
      row();
      nvortex = atoi_f(v[0]);
      for (int i = 0; i < nvortex; i++) row();
    } else if (key == "VORTEXPAIR_CROW" || key == "VORTEXPAIR_SECTION") {
      for (int i = 0; i < nvortex; i++) row();
    } else if (key == "POSTPROCESS") {
      row();
      const int n = atoi_f(v[0]);
      for (int i = 0; i < n; i++) row();
    } else if (key == "END") {
      break;
    } else {
      static const char *one[] = {"DIMENSION", "CHOPS", "ELL", "ZLEN", "MINC", "MKLINK", "DT", "VISCOSITY",
                                  "BOUSSINESQ", "VORTEXPAIR_NOISE", "TIME_INTEG", "FREESTREAM", "STRAIN", "LINEAR",
                                  "DEBUG"};
      for (const char *k : one)
        if (key == k) {
          row();
          break;
        }
    }
  }
  return 0;
}
