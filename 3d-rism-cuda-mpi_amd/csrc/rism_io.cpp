// rism_io.cpp -- the reference's file formats, host only (no device use).
//
//   .inp  control + solute     read_input.cu:21-97, solute.cu:18-54
//   .xvk  solvent chi_vv(k)     solvent_read.cc:10-94
//   .xmu  thermodynamics        output_xmu.cu:7-82
//
// Errors are returned as RISM_ERR_* codes with a message retrievable through
// rism_last_error(); the reference prints and calls exit(1) instead.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/rism3d_c.h"
#include "rism_internal.h"

namespace {

const double kAvogadro = 6.02214076e23;  // physical.h:5

bool is_natural_number(const std::string &s) {  // read_input.cu:8-19
  try {
    size_t pos;
    int num = std::stoi(s, &pos);
    if (pos != s.length()) return false;
    return num >= 0;
  } catch (...) {
    return false;
  }
}

double *copy_array(const std::vector<double> &v) {
  double *p = (double *)std::malloc(sizeof(double) * (v.empty() ? 1 : v.size()));
  if (!v.empty()) std::memcpy(p, v.data(), sizeof(double) * v.size());
  return p;
}

// solute.cu:18-48: integer-Angstrom shift centring the bounding box.
void centre(std::vector<double> &xyz, int natu, double shift[3]) {
  double mn[3], mx[3];
  for (int d = 0; d < 3; ++d) {
    mn[d] = std::numeric_limits<double>::max();
    mx[d] = std::numeric_limits<double>::lowest();
  }
  for (int n = 0; n < natu; ++n)
    for (int d = 0; d < 3; ++d) {
      mn[d] = std::min(mn[d], xyz[3 * n + d]);
      mx[d] = std::max(mx[d], xyz[3 * n + d]);
    }
  for (int d = 0; d < 3; ++d) shift[d] = std::round(-(mx[d] - mn[d]) / 2 - mn[d]);
  for (int n = 0; n < natu; ++n)
    for (int d = 0; d < 3; ++d) xyz[3 * n + d] += shift[d];
}

int read_xvk_into(const std::string &path, rism_input *in) {
  std::ifstream f(path);
  if (!f) return rism::set_error(RISM_ERR_IO, "Solvent file could not be opened! (" + path + ")");
  std::vector<std::string> lines;
  std::string line;
  // header: title, "## num natv ntab dr dk", "## _ T xt", column names,
  // num site lines, a "##" line; then ntab records of natv^2 reals
  for (int i = 0; i < 3 && std::getline(f, line); ++i) lines.push_back(line);
  if (lines.size() < 3) return rism::set_error(RISM_ERR_IO, "truncated .xvk header: " + path);
  std::string dummy;
  int num = 0, natv = 0, ntab = 0;
  double dr = 0;
  {
    std::istringstream iss(lines[1]);
    iss >> dummy >> num >> natv >> ntab >> dr;
    if (!iss || num <= 0 || natv <= 0 || ntab < 3)
      return rism::set_error(RISM_ERR_IO, "bad .xvk header line: " + lines[1]);
  }
  if (natv > RISM_MAX_SITES)
    return rism::set_error(RISM_ERR_ARG, "more solvent sites than supported (16)");
  double temper = 0, xt = 0;
  {
    std::istringstream iss(lines[2]);
    iss >> dummy >> dummy >> temper >> xt;
    if (!iss) return rism::set_error(RISM_ERR_IO, "bad .xvk header line: " + lines[2]);
  }
  std::getline(f, line);  // column names
  std::vector<double> sig(natv, 0.0), eps(natv, 0.0), q(natv, 0.0), rho(natv, 0.0);
  for (int i = 0; i < num; ++i) {
    if (!std::getline(f, line)) return rism::set_error(RISM_ERR_IO, "truncated .xvk sites");
    std::istringstream iss(line);
    int sym;
    double s, e, c, den;
    iss >> dummy >> dummy >> dummy >> sym >> s >> e >> c >> dummy >> dummy >> dummy >> den;
    if (!iss) return rism::set_error(RISM_ERR_IO, "bad .xvk site line: " + line);
    int i2 = std::abs(sym) - 1;
    if (i2 < 0 || i2 >= natv) return rism::set_error(RISM_ERR_IO, "bad site symmetry id");
    sig[i2] = s;  // the last listed member of a merged site wins
    eps[i2] = e;
    q[i2] = c;
    rho[i2] += den;  // densities of merged sites add
  }
  for (int i = 0; i < natv; ++i) rho[i] *= kAvogadro * 1.0e-27;
  std::getline(f, line);  // "##"
  const size_t need = (size_t)ntab * natv * natv;
  std::vector<double> xvv(need);
  for (size_t i = 0; i < need; ++i) {
    std::string tok;
    if (!(f >> tok)) return rism::set_error(RISM_ERR_IO, "truncated .xvk table: " + path);
    xvv[i] = std::strtod(tok.c_str(), nullptr);
  }
  in->natv = natv;
  in->ntab = ntab;
  in->dr = dr;
  in->temper = temper;
  in->xt = xt;
  in->vsig = copy_array(sig);
  in->veps = copy_array(eps);
  in->vq = copy_array(q);
  in->vrho = copy_array(rho);
  in->xvv = copy_array(xvv);
  return RISM_OK;
}

}  // namespace

extern "C" {

const char *rism_version(void) { return "20220301"; }

void rism_free_input(rism_input *in) {
  if (!in) return;
  std::free(in->usig);
  std::free(in->ueps);
  std::free(in->uq);
  std::free(in->uxyz);
  std::free(in->vsig);
  std::free(in->veps);
  std::free(in->vq);
  std::free(in->vrho);
  std::free(in->xvv);
  std::free(in);
}

int rism_read_xvk(const char *path, rism_input **out) {
  if (!path || !out) return rism::set_error(RISM_ERR_ARG, "null argument");
  rism_input *in = (rism_input *)std::calloc(1, sizeof(rism_input));
  int rc = read_xvk_into(path, in);
  if (rc != RISM_OK) {
    rism_free_input(in);
    return rc;
  }
  *out = in;
  return RISM_OK;
}

int rism_read_input(const char *control, const char *structure, int center, int q0,
                    const char *solvent_dir, rism_input **out) {
  if (!control || !out) return rism::set_error(RISM_ERR_ARG, "null argument");
  std::ifstream f(control);
  if (!f) return rism::set_error(RISM_ERR_IO, std::string("cannot open ") + control);
  rism_input *in = (rism_input *)std::calloc(1, sizeof(rism_input));
  auto fail = [&](int code, const std::string &msg) {
    rism_free_input(in);
    return rism::set_error(code, msg);
  };
  std::string outlist, check, closure, fsolvent;
  f >> outlist >> in->ksave >> check;
  if (!f) return fail(RISM_ERR_IO, "cannot parse line 1 of the input file");
  if (check != "20220301") return fail(RISM_ERR_VERSION, "Error: This input file is for old version.");
  f >> closure;
  if (closure == "KH") {
    in->closure = RISM_CLOSURE_KH;
  } else if (closure == "HNC") {
    in->closure = RISM_CLOSURE_HNC;
  } else {
    // PSEn / PSE-n (not in the reference's reader)
    int n = 0;
    if (closure.size() > 3 && closure.compare(0, 3, "PSE") == 0) {
      const std::string num = closure.substr(closure[3] == '-' ? 4 : 3);
      if (!num.empty() && num.find_first_not_of("0123456789") == std::string::npos &&
          num.size() <= 2)
        n = std::atoi(num.c_str());
    }
    if (n < 1 || n > 20) return fail(RISM_ERR_CLOSURE, "Error: Unexpected closure switch ");
    in->closure = RISM_CLOSURE_PSE(n);
  }
  f >> fsolvent;
  f >> in->convergence >> in->maxstep;
  f >> in->an2_count >> in->an2_m >> in->an2_mp;
  f >> in->box[0] >> in->box[1] >> in->box[2];
  f >> in->grid[0] >> in->grid[1] >> in->grid[2];
  if (!f) return fail(RISM_ERR_IO, "cannot parse the control lines of the input file");
  std::snprintf(in->outlist, sizeof(in->outlist), "%s", outlist.c_str());
  std::snprintf(in->fsolvent, sizeof(in->fsolvent), "%s", fsolvent.c_str());

  std::ifstream fs;
  std::istream *src = &f;
  if (structure && structure[0]) {
    fs.open(structure);
    if (!fs) return fail(RISM_ERR_IO, std::string("cannot open ") + structure);
    src = &fs;
  }
  std::string tmp;
  *src >> tmp;
  if (!is_natural_number(tmp))
    return fail(RISM_ERR_NATOM, "Error: Number of solute atom is not a natural number.");
  const int natu = std::stoi(tmp);
  std::vector<double> sig(natu), eps(natu), q(natu), xyz(3 * (size_t)natu);
  for (int i = 0; i < natu; ++i) {
    *src >> sig[i] >> eps[i] >> q[i] >> xyz[3 * i] >> xyz[3 * i + 1] >> xyz[3 * i + 2];
    if (!*src) return fail(RISM_ERR_IO, "truncated solute atom list");
  }
  in->shift[0] = in->shift[1] = in->shift[2] = 0.0;
  if (center) centre(xyz, natu, in->shift);
  if (q0) std::fill(q.begin(), q.end(), 0.0);  // solute.cu:50-54
  in->natu = natu;
  in->usig = copy_array(sig);
  in->ueps = copy_array(eps);
  in->uq = copy_array(q);
  in->uxyz = copy_array(xyz);

  std::string spath = fsolvent;
  if (solvent_dir && solvent_dir[0] && fsolvent[0] != '/')
    spath = std::string(solvent_dir) + "/" + fsolvent;
  int rc = read_xvk_into(spath, in);
  if (rc != RISM_OK) {
    rism_free_input(in);
    return rc;
  }
  *out = in;
  return RISM_OK;
}

int rism_write_xmu(const char *path, int closure, const rism_thermo_out *th) {
  if (!path || !th) return rism::set_error(RISM_ERR_ARG, "null argument");
  std::ofstream o(path);
  if (!o) return rism::set_error(RISM_ERR_IO, std::string("cannot write ") + path);
  const int n = th->natv;
  o << " $RESULTS" << std::endl;
  o << "SFE_SC= " << std::fixed << std::setprecision(5) << th->sfe_sc << " !(J/mol)" << std::endl;
  for (int v = 0; v < n; ++v)
    o << "  SFEC_SC(" << v << ")= " << std::fixed << std::setprecision(5) << th->sfec_sc[v]
      << std::endl;
  o << std::endl;
  if (closure == RISM_CLOSURE_KH) {
    o << "SFE_SC_HNC= " << std::fixed << std::setprecision(5) << th->sfe_sc_hnc << " !(J/mol)"
      << std::endl;
    for (int v = 0; v < n; ++v)
      o << "  SFEC_SC_HNC(" << v << ")= " << std::fixed << std::setprecision(5)
        << th->sfec_sc_hnc[v] << std::endl;
    o << std::endl;
  }
  o << "SFE_GF= " << std::fixed << std::setprecision(5) << th->sfe_gf << " !(J/mol)" << std::endl;
  for (int v = 0; v < n; ++v)
    o << "  SFEC_GF(" << v << ")= " << std::fixed << std::setprecision(5) << th->sfec_gf[v]
      << std::endl;
  o << std::endl;
  o << "PMV= " << std::fixed << std::setprecision(5) << th->pmv << " !(L/mol)" << std::endl;
  o << "Pressure= " << std::fixed << std::setprecision(5) << th->pressure << " !(J/m^3)"
    << std::endl;
  o << "Correction_Term= " << std::fixed << std::setprecision(5) << th->correction
    << " !(J/mol)" << std::endl;
  o << " $END" << std::endl;
  return o ? RISM_OK : rism::set_error(RISM_ERR_IO, std::string("write failed: ") + path);
}

}  // extern "C"
