// rism3d -- command-line drop-in for one solve of the reference executable
// (main.cu:6-69): same getopt string "c:i:s:l:e:f:z", same .inp/.xvk inputs,
// same stdout progress lines, same .xmu/.xmu0/.guv/.tuv outputs, on one GPU
// through the C ABI of librism3d.
//
//   rism3d [-i control] [-s solute] [-l lambda | -e lambda] [-f x] [-z]
//          [-c dev] [input.inp]
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/rism3d_c.h"

namespace {

int die(const char *what) {
  std::cout << what << ": " << rism_last_error() << std::endl;
  return 1;
}

bool file_exists(const std::string &p) {
  std::ifstream f(p);
  return f.is_open();
}

}  // namespace

int main(int argc, char *argv[]) {
  std::string input, structure;
  bool centering = true, q0 = false;
  int adswitch = 0;
  double lambda = 1.0;
  int ch;
  while ((ch = getopt(argc, argv, "c:i:s:l:e:f:z")) != -1) {
    switch (ch) {
      case 'c':  // parsed and unused, as in the reference (main.cu:18-20)
        break;
      case 'i':
        input = optarg;
        break;
      case 's':
        structure = optarg;
        break;
      case 'l':  // set_ad(lambda, 1)
        lambda = std::atof(optarg);
        adswitch = 1;
        break;
      case 'e':  // set_ad(lambda, 2)
        lambda = std::atof(optarg);
        adswitch = 2;
        break;
      case 'f':
        centering = false;
        break;
      case 'z':
        q0 = true;
        break;
    }
  }
  if (input.empty() || structure.empty()) {  // main.cu:42-48
    if (optind >= argc || argv[optind] == nullptr) {
      std::cout << "No input file!" << std::endl;
      return 1;
    }
    input = argv[optind];
  }

  std::cout << "reading input data file  : " << input << std::endl;
  if (!structure.empty()) std::cout << "reading solute data file : " << structure << std::endl;
  rism_input *in = nullptr;
  int rc = rism_read_input(input.c_str(), structure.empty() ? nullptr : structure.c_str(),
                           centering, q0, nullptr, &in);
  if (rc != RISM_OK) {
    std::cout << rism_last_error() << std::endl;
    return 1;
  }
  // cell.cc:20-25
  const double maxdr = 0.5;
  if (in->box[0] / in->grid[0] > maxdr || in->box[1] / in->grid[1] > maxdr ||
      in->box[2] / in->grid[2] > maxdr) {
    std::cout << "##########################################" << std::endl;
    std::cout << "WARRING: Grid spacing is greater than " << maxdr << "." << std::endl;
    std::cout << "##########################################" << std::endl;
  }

  rism_grid g{in->grid[0], in->grid[1], in->grid[2], {in->box[0], in->box[1], in->box[2]}};
  rism_solvent sv{in->natv, in->ntab, in->dr,   in->temper, in->xt,
                  in->vsig, in->veps, in->vq,   in->vrho,   in->xvv};
  rism_solute su{in->natu, in->usig, in->ueps, in->uq, in->uxyz};
  rism_opts op{};
  op.closure = in->closure;
  op.an2_count = in->an2_count;
  op.an2_m = in->an2_m;
  op.an2_mp = in->an2_mp;
  op.ad_switch = adswitch;
  op.lambda = lambda;
  op.verbose = 1;
  int err = 0;
  rism_ctx *ctx = rism_create(&g, &sv, &su, &op, &err);
  if (!ctx) {
    std::cout << rism_last_error() << std::endl;
    rism_free_input(in);
    return 1;
  }
  // set_fname.cu:5-12
  std::string fname = structure.empty() ? input : structure;
  fname = fname.substr(0, fname.rfind("."));

  if ((rc = rism_setup_potentials(ctx)) != RISM_OK) return die("potentials");
  const std::string ftuv = fname + ".tuv";
  if (file_exists(ftuv)) {  // iterate.cu:26-35, with the restored t uploaded
    if ((rc = rism_read_tuv(ctx, ftuv.c_str())) != RISM_OK) return die("read_tuv");
  }

  // iterate.cu:45-82, with the .tuv checkpoints of ksave
  std::cout << "relaxing 3D UV RISM:" << std::endl;
  bool conver = false, diverge = false;
  for (int istep = 1; istep <= in->maxstep; ++istep) {
    double rms = 0.0;
    if ((rc = rism_calculate(ctx, &rms)) != RISM_OK) return die("calculate");
    diverge = !std::isfinite(rms);
    if (diverge) break;
    if (rms <= in->convergence)
      conver = true;
    else if ((rc = rism_anderson(ctx)) != RISM_OK)
      return die("anderson");
    std::cout << " Step = " << istep << " Reside = " << rms << std::endl;
    if (in->ksave > 0 && istep % in->ksave == 0) {
      std::cout << "saving Tuv to file:  " << ftuv << "  ..." << std::endl;
      if ((rc = rism_write_tuv(ctx, ftuv.c_str())) != RISM_OK) return die("write_tuv");
      std::cout << "done." << std::endl;
    }
    if (conver) {
      if (in->ksave != 0) {
        std::cout << "saving Tuv to file:  " << ftuv << "  ..." << std::endl;
        if ((rc = rism_write_tuv(ctx, ftuv.c_str())) != RISM_OK) return die("write_tuv");
        std::cout << "done." << std::endl;
      }
      break;
    }
  }
  if (diverge)
    std::cout << "Calculation diverged." << std::endl;
  else if (!conver)
    std::cout << "3D UV RISM: reached limit # of relaxation steps: " << in->maxstep << std::endl;

  // output.cu:6-106
  std::string outlist = in->outlist;
  std::transform(outlist.begin(), outlist.end(), outlist.begin(), ::tolower);
  if (outlist.find('m') != std::string::npos) {
    rism_thermo_out th;
    if ((rc = rism_thermo(ctx, &th)) != RISM_OK) return die("thermo");
    const std::string fxmu = fname + (q0 ? ".xmu0" : ".xmu");
    if ((rc = rism_write_xmu(fxmu.c_str(), in->closure, &th)) != RISM_OK) return die("write_xmu");
  }
  if (outlist.find('d') != std::string::npos) {  // cal_grad + output_grad.cu:8-22
    const std::string f = fname + ".gra";
    std::vector<double> du((size_t)in->natu * 3);
    if ((rc = rism_analysis(ctx, RISM_ANALYSIS_GRAD, du.data())) != RISM_OK) return die("grad");
    std::cout << "outputting grad to file:  " << f << "  ..." << std::endl;
    std::ofstream o(f);
    for (int a = 0; a < in->natu; ++a)
      o << du[3 * a] << " " << du[3 * a + 1] << " " << du[3 * a + 2] << std::endl;
    std::cout << "done." << std::endl;
  }
  if (outlist.find('c') != std::string::npos) {  // output_cuv.cu:9-17 (rank 0)
    const std::string f = fname + ".cuv";
    std::cout << "outputting Cuv to file:  " << f << "  ..." << std::endl;
    if ((rc = rism_write_cuv(ctx, (f + "0").c_str())) != RISM_OK) return die("write_cuv");
    std::cout << "done." << std::endl;
  }
  if (outlist.find('g') != std::string::npos) {
    const std::string fguv = fname + ".guv";
    std::cout << "outputting Guv to file:  " << fguv << "  ..." << std::endl;
    if ((rc = rism_write_guv(ctx, fguv.c_str(), in->shift)) != RISM_OK) return die("write_guv");
    std::cout << "done." << std::endl;
  }
  if (outlist.find('h') != std::string::npos) {  // output_huv.cu:9-17 (rank 0)
    const std::string f = fname + ".huv";
    std::cout << "outputting Huv to file:  " << f << "  ..." << std::endl;
    if ((rc = rism_write_huv(ctx, (f + "0").c_str())) != RISM_OK) return die("write_huv");
    std::cout << "done." << std::endl;
  }
  if (outlist.find('a') != std::string::npos) {  // cal_ad1/cal_ad2 + output_ad.cu:8-22
    std::stringstream ss;
    ss << fname << "-" << adswitch << "-" << lambda;
    std::vector<double> du(in->natu);
    if ((rc = rism_analysis(ctx, RISM_ANALYSIS_AD, du.data())) != RISM_OK) return die("ad");
    std::cout << "outputting ad to file:  " << ss.str() << "  ..." << std::endl;
    std::ofstream o(ss.str());
    for (int a = 0; a < in->natu; ++a) o << du[a] << std::endl;
    std::cout << "done." << std::endl;
  }
  if (outlist.find('b') != std::string::npos) {  // cal_euv + output_euv.cu:8-31
    const std::string f = fname + ".euv";
    const int w = 1 + in->natv;
    std::vector<double> e((size_t)in->natu * w);
    if ((rc = rism_analysis(ctx, RISM_ANALYSIS_EUV, e.data())) != RISM_OK) return die("euv");
    std::cout << "outputting euv to file:  " << f << "  ..." << std::endl;
    std::ofstream o(f);
    for (int a = 0; a < in->natu; ++a) {
      o << std::fixed << e[(size_t)a * w];
      for (int v = 1; v < w; ++v) o << " " << std::fixed << e[(size_t)a * w + v];
      o << std::endl;
    }
    std::cout << "done." << std::endl;
  }
  rism_destroy(ctx);
  rism_free_input(in);
  return 0;
}
