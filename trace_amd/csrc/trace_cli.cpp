// Reconstruction programs: HDF5 in, iterations on the GPU, HDF5 out.
//
// One source, five programs (TRACE_CLI_APP):
//   0 sirt   sirt_main.cc       1 mlem   mlem_main.cc
//   2 pml    pml_main.cc        3 apmlr  apmlr_main.cc
//   4 trace  src/all/trace.cc (the draft unified program: long options,
//            `-r algorithm`, one input file holding both datasets)
// Each takes the reference program's flags (same short and long names,
// defaults and required set), echoes its configuration the same way,
// applies the same input preprocessing and initial image, and writes the
// [S][N][N] result where the reference would.  The engine loop of the
// reference (RunParallelReduction / ParInPlaceLocalSynch / UpdateRecon /
// ResetReductionSpaces) is one trace_reconstruct call per iteration.
//
// Command-line errors follow the reference's TCLAP behaviour
// (CmdLine.h:485-508, StdOutput.h:131-153): "PARSE ERROR:" on stderr, a
// usage hint, exit status 1.  Input or GPU failures print "error: ..." and
// exit with status 2 (the reference throws and aborts).
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "trace_h5io.h"
#include "trace_hip.h"

#ifndef TRACE_CLI_APP
#error "TRACE_CLI_APP selects the program (0 sirt, 1 mlem, 2 pml, 3 apmlr, 4 trace)"
#endif

namespace {

enum App { kSirt = 0, kMlem = 1, kPml = 2, kApmlr = 3, kUnified = 4 };
constexpr App kApp = static_cast<App>(TRACE_CLI_APP);

enum Kind { kStr, kInt, kFloat, kSwitch };

struct Opt {
  const char *sname;  // short name without '-', or "" for long-only
  const char *lname;  // long name without "--"
  Kind kind;
  bool required;
  const char *def;
  const char *desc;
  std::string val;
  bool set = false;
};

struct ParseError {
  std::string id, msg;
};

class CmdLine {
 public:
  CmdLine(const char *prog, const char *title) : prog_(prog), title_(title) {}
  void add(const char *s, const char *l, Kind k, bool req, const char *def, const char *desc) {
    opts_.push_back(Opt{s, l, k, req, def, desc, def ? def : "", false});
  }
  Opt &get(const char *lname) {
    for (auto &o : opts_)
      if (!strcmp(o.lname, lname)) return o;
    std::fprintf(stderr, "internal: no option %s\n", lname);
    std::exit(3);
  }
  std::string str(const char *l) { return get(l).val; }
  int integer(const char *l) { return (int)std::strtol(get(l).val.c_str(), nullptr, 10); }
  float real(const char *l) { return std::strtof(get(l).val.c_str(), nullptr); }
  bool on(const char *l) { return get(l).set; }
  void allowed(const char *l, std::vector<std::string> v) {
    allowed_lname_ = l;
    allowed_ = std::move(v);
  }

  void parse(int argc, char **argv) {
    try {
      for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "-h" || a == "--help") {
          usage(std::cout);
          std::exit(0);
        }
        if (a == "--version") {
          std::cout << prog_ << "  version: 0.01 (MI355X)\n";
          std::exit(0);
        }
        Opt *o = nullptr;
        std::string name, value;
        bool has_value = false;
        const size_t eq = a.find('=');
        if (a.rfind("--", 0) == 0 && a.size() > 2) {
          name = a.substr(2, eq == std::string::npos ? std::string::npos : eq - 2);
          for (auto &c : opts_)
            if (name == c.lname) o = &c;
        } else if (a.size() > 1 && a[0] == '-') {
          name = a.substr(1, eq == std::string::npos ? std::string::npos : eq - 1);
          for (auto &c : opts_)
            if (*c.sname && name == c.sname) o = &c;
        }
        if (!o) throw ParseError{"undefined", "Couldn't find match for argument " + a};
        if (eq != std::string::npos) {
          value = a.substr(eq + 1);
          has_value = true;
        }
        const std::string id = label(*o);
        if (o->set) throw ParseError{id, "Argument already set!"};
        if (o->kind == kSwitch) {
          if (has_value) throw ParseError{id, "Switch takes no value: " + a};
          o->set = true;
          continue;
        }
        if (!has_value) {
          if (i + 1 >= argc) throw ParseError{id, "Missing a value for this argument!"};
          value = argv[++i];
        }
        check_value(*o, value);
        o->val = value;
        o->set = true;
      }
      std::string missing;
      for (auto &o : opts_)
        if (o.required && !o.set) missing += (missing.empty() ? "" : ", ") + std::string(o.lname);
      if (!missing.empty())
        throw ParseError{"undefined", "Required argument(s) missing: " + missing};
    } catch (const ParseError &e) {
      std::cerr << "PARSE ERROR: " << e.id << "\n             " << e.msg << "\n\n"
                << "For complete USAGE and HELP type: \n   " << prog_ << " --help\n\n";
      std::exit(1);
    }
  }

 private:
  static std::string label(const Opt &o) {
    std::string s = *o.sname ? std::string("-") + o.sname + " (--" + o.lname + ")"
                             : std::string("--") + o.lname;
    return "Argument: " + s;
  }
  void check_value(const Opt &o, const std::string &v) {
    char *end = nullptr;
    if (o.kind == kInt) {
      std::strtol(v.c_str(), &end, 10);
      if (v.empty() || *end)
        throw ParseError{label(o), "Couldn't read argument value from string '" + v + "'"};
    } else if (o.kind == kFloat) {
      std::strtof(v.c_str(), &end);
      if (v.empty() || *end)
        throw ParseError{label(o), "Couldn't read argument value from string '" + v + "'"};
    }
    if (!allowed_.empty() && !strcmp(o.lname, allowed_lname_)) {
      for (auto &a : allowed_)
        if (a == v) return;
      std::string c;
      for (auto &a : allowed_) c += (c.empty() ? "" : "|") + a;
      throw ParseError{label(o), "Value '" + v + "' does not meet constraint: " + c};
    }
  }
  void usage(std::ostream &os) {
    os << "\nUSAGE: \n\n   " << prog_;
    for (auto &o : opts_) {
      const std::string f = *o.sname ? std::string("-") + o.sname : std::string("--") + o.lname;
      os << " " << (o.required ? "" : "[") << f
         << (o.kind == kSwitch ? "" : (o.kind == kStr ? " <string>" : o.kind == kInt ? " <int>" : " <float>"))
         << (o.required ? "" : "]");
    }
    os << "\n\nWhere: \n\n";
    for (auto &o : opts_) {
      os << "   ";
      if (*o.sname) os << "-" << o.sname << ",  ";
      os << "--" << o.lname << "\n     " << (o.required ? "(required)  " : "") << o.desc;
      if (!o.required && o.def && *o.def && o.kind != kSwitch) os << " (default " << o.def << ")";
      os << "\n\n";
    }
    os << "   -h,  --help\n     Displays usage information and exits.\n\n   --version\n"
       << "     Displays version information and exits.\n\n   " << title_ << "\n\n";
  }

  const char *prog_, *title_;
  std::vector<Opt> opts_;
  const char *allowed_lname_ = "";
  std::vector<std::string> allowed_;
};

[[noreturn]] void die(const std::string &msg) {
  std::cerr << "error: " << msg << std::endl;
  std::exit(2);
}

void h5_check(int rc, const char *what) {
  if (rc != TRACE_H5_OK) die(std::string(what) + ": " + trace_h5_last_error());
}

void trace_check(int rc, const char *what) {
  if (rc != TRACE_OK) die(std::string(what) + ": " + trace_last_error());
}

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) die(std::string(what) + ": " + hipGetErrorString(e));
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

struct Config {
  trace_alg_t alg;
  std::string proj_file, proj_dset, theta_file, theta_dset, out_file, out_dset;
  int iters = 0, threads = 1, device = 0;
  float center = 0.f;
  bool deg2rad = true;
  trace_params_t prm{};
};

void common_main_flags(CmdLine &cmd) {
  // sirt_main.cc:25-46 (identical in mlem_main.cc, pml_main.cc, apmlr_main.cc)
  cmd.add("f", "projectionFilePath", kStr, true, "", "Projection file path");
  cmd.add("d", "projectionDatasetPath", kStr, false, "/exchange/data",
          "Projection dataset path in hdf5");
  cmd.add("e", "thetaFilePath", kStr, true, "", "Theta file path");
  cmd.add("q", "thetaDatasetPath", kStr, false, "/exchange/theta", "Theta dataset path");
  cmd.add("o", "reconOutputPath", kStr, false, "./output.h5",
          "Output file path for reconstructed image (hdf5)");
  cmd.add("r", "reconDatasetPath", kStr, false, "/data",
          "Reconstruction dataset path in hdf5 file");
  cmd.add("i", "iteration", kInt, true, "0", "Number of iterations");
  cmd.add("c", "center", kFloat, false, "0", "Center value");
  cmd.add("t", "thread", kInt, false, "1",
          "Number of threads per process (accepted; the GPU path ignores it)");
}

Config parse_main(int argc, char **argv) {
  static const char *names[] = {"sirt", "mlem", "pml", "apmlr"};
  static const char *titles[] = {"SIRT Iterative Image Reconstruction",
                                 "MLEM Iterative Image Reconstruction",
                                 "PML Iterative Image Reconstruction",
                                 "APMLR Iterative Image Reconstruction"};
  CmdLine cmd(names[kApp], titles[kApp]);
  common_main_flags(cmd);
  if (kApp == kPml) cmd.add("j", "beta", kFloat, false, "10", "Beta value");  // pml_main.cc:48-49
  if (kApp == kApmlr) {  // apmlr_main.cc:54-63
    cmd.add("j", "beta0", kFloat, false, "10", "Beta 0 value");
    cmd.add("k", "beta1", kFloat, false, "1", "Beta 1 value");
    cmd.add("l", "delta0", kFloat, false, "1", "Delta 0 value");
    cmd.add("m", "delta1", kFloat, false, "1", "Delta 1 value");
    cmd.add("w", "regw", kFloat, false, "1", "Regw value");
    cmd.add("", "apmlr-fixed", kSwitch, false, "",
            "Project slice s with slice s (the reference projects it with slice s-1); "
            "no counterpart in the reference");
  }
  cmd.add("", "device", kInt, false, "0", "GPU ordinal; no counterpart in the reference");
  cmd.parse(argc, argv);

  Config c;
  c.alg = static_cast<trace_alg_t>(kApp);
  c.proj_file = cmd.str("projectionFilePath");
  c.proj_dset = cmd.str("projectionDatasetPath");
  c.theta_file = cmd.str("thetaFilePath");
  c.theta_dset = cmd.str("thetaDatasetPath");
  c.out_file = cmd.str("reconOutputPath");
  c.out_dset = cmd.str("reconDatasetPath");
  c.iters = cmd.integer("iteration");
  c.center = cmd.real("center");
  c.threads = cmd.integer("thread");
  c.device = cmd.integer("device");
  c.deg2rad = kApp != kPml;  // pml_main.cc:121 leaves DegreeToRadian commented out
  c.prm.beta = 10.f;
  c.prm.beta1 = 1.f;
  c.prm.delta = 1.f;
  c.prm.delta1 = 1.f;
  c.prm.regw = 1.f;
  c.prm.apmlr_compat = 1;
  if (kApp == kPml) c.prm.beta = cmd.real("beta");
  if (kApp == kApmlr) {
    c.prm.beta = cmd.real("beta0");
    c.prm.beta1 = cmd.real("beta1");
    c.prm.delta = cmd.real("delta0");
    c.prm.delta1 = cmd.real("delta1");
    c.prm.regw = cmd.real("regw");
    c.prm.apmlr_compat = cmd.on("apmlr-fixed") ? 0 : 1;
  }

  // configuration echo (sirt_main.cc:69-81, pml_main.cc:87, apmlr_main.cc:109-113)
  std::cout << "MPI rank:0; MPI size:1" << std::endl;
  std::cout << "Projection file path=" << c.proj_file << std::endl;
  std::cout << "Projection dataset path in hdf5=" << c.proj_dset << std::endl;
  std::cout << "Theta file path=" << c.theta_file << std::endl;
  std::cout << "Theta dataset path=" << c.theta_dset << std::endl;
  std::cout << "Output file path=" << c.out_file << std::endl;
  std::cout << "Recon. dataset path=" << c.out_dset << std::endl;
  std::cout << "Number of iterations=" << c.iters << std::endl;
  std::cout << "Center value=" << c.center << std::endl;
  std::cout << "Number of threads per process=" << c.threads << std::endl;
  if (kApp == kPml) std::cout << "Beta=" << c.prm.beta << std::endl;
  if (kApp == kApmlr) {
    std::cout << "Beta0=" << c.prm.beta << std::endl;
    std::cout << "Beta1=" << c.prm.beta1 << std::endl;
    std::cout << "Delta0=" << c.prm.delta << std::endl;
    std::cout << "Delta1=" << c.prm.delta1 << std::endl;
    std::cout << "Regw=" << c.prm.regw << std::endl;
  }
  return c;
}

Config parse_unified(int argc, char **argv) {
  // src/all/trace.cc:35-150
  CmdLine cmd("trace", "Trace: Iterative Image Reconstruction Engine");
  cmd.add("", "degree-to-radian", kSwitch, false, "", "Transform theta values from degree to radian");
  cmd.add("", "regw", kFloat, false, "1", "Regw 1 value.");
  cmd.add("", "beta-one", kFloat, false, "1", "Beta 1 value.");
  cmd.add("", "beta-zero", kFloat, false, "10", "Beta 0 value.");
  cmd.add("", "delta-one", kFloat, false, "1", "Delta 1 value.");
  cmd.add("", "delta-zero", kFloat, false, "1", "Delta 0 value.");
  cmd.add("p", "projection-file-path", kStr, false, "", "Input projection file path.");
  cmd.add("d", "projection-dataset-path", kStr, false, "/exchange/data",
          "Projection dataset path in HDF5 file.");
  cmd.add("e", "theta-file-path", kStr, false, "", "Input theta file path.");
  cmd.add("q", "theta-dataset-path", kStr, false, "/exchange/theta",
          "Theta dataset path in HDF5 file.");
  cmd.add("m", "output-dataset-path", kStr, false, "/image",
          "Reconstructed 3D image dataset path in hdf5 file.");
  cmd.add("c", "center", kFloat, false, "0", "Center value for input dataset.");
  cmd.add("t", "thread-count", kInt, false, "1",
          "Number of threads per process (accepted; the GPU path ignores it).");
  cmd.add("i", "iteration-count", kInt, true, "0", "Number of iterations.");
  cmd.add("o", "output-file-path", kStr, false, "./output.h5",
          "Output HDF5 file path.");
  cmd.add("f", "input-file-path", kStr, false, "",
          "Input HDF5 file path. This file should include both projection and theta datasets.");
  cmd.add("r", "reconstruction-algorithm", kStr, true, "",
          "Reconstruction algorithm to be applied.");
  cmd.allowed("reconstruction-algorithm", {"sirt", "mlem", "pml", "apmlr"});
  cmd.add("", "apmlr-fixed", kSwitch, false, "",
          "APMLR: project slice s with slice s (the reference projects it with slice s-1); "
          "no counterpart in the reference");
  cmd.add("", "device", kInt, false, "0", "GPU ordinal; no counterpart in the reference");
  cmd.parse(argc, argv);

  // input file rules (trace.cc:123-141); the reference exits with status 0
  // here, this program with the parse-error status
  const bool sep = cmd.on("projection-file-path") || cmd.on("theta-file-path");
  if (sep == cmd.on("input-file-path")) {
    std::cerr << "Both input and projection (or theta) files are set (or unset). Only one of "
                 "them should be given."
              << std::endl;
    std::exit(1);
  }
  Config c;
  if (sep) {
    if (!(cmd.on("projection-file-path") && cmd.on("theta-file-path"))) {
      std::cerr << "Both projection and theta files need to be set." << std::endl;
      std::exit(1);
    }
    c.proj_file = cmd.str("projection-file-path");
    c.theta_file = cmd.str("theta-file-path");
  } else {
    c.proj_file = c.theta_file = cmd.str("input-file-path");
  }
  const std::string a = cmd.str("reconstruction-algorithm");
  c.alg = a == "sirt" ? TRACE_SIRT : a == "mlem" ? TRACE_MLEM : a == "pml" ? TRACE_PML : TRACE_APMLR;
  c.proj_dset = cmd.str("projection-dataset-path");
  c.theta_dset = cmd.str("theta-dataset-path");
  c.out_file = cmd.str("output-file-path");
  c.out_dset = cmd.str("output-dataset-path");
  c.iters = cmd.integer("iteration-count");
  c.center = cmd.real("center");
  c.threads = cmd.integer("thread-count");
  c.device = cmd.integer("device");
  c.deg2rad = cmd.on("degree-to-radian");
  c.prm.beta = cmd.real("beta-zero");
  c.prm.beta1 = cmd.real("beta-one");
  c.prm.delta = cmd.real("delta-zero");
  c.prm.delta1 = cmd.real("delta-one");
  c.prm.regw = cmd.real("regw");
  c.prm.apmlr_compat = cmd.on("apmlr-fixed") ? 0 : 1;

  // configuration echo (trace.cc:151-178)
  std::cout << "Reconstruction algorithm=" << a << std::endl;
  std::cout << "Projection file path=" << c.proj_file << std::endl;
  std::cout << "Projection dataset path in hdf5=" << c.proj_dset << std::endl;
  std::cout << "Theta file path=" << c.theta_file << std::endl;
  std::cout << "Theta dataset path=" << c.theta_dset << std::endl;
  std::cout << "Output file path=" << c.out_file << std::endl;
  std::cout << "Recon. dataset path=" << c.out_dset << std::endl;
  std::cout << "Number of iterations=" << c.iters << std::endl;
  std::cout << "Center value=" << c.center << std::endl;
  std::cout << "Number of threads per process=" << c.threads << std::endl;
  std::cout << "Degree to radian=" << c.deg2rad << std::endl;
  if (c.alg == TRACE_PML || c.alg == TRACE_APMLR) std::cout << "Beta 0=" << c.prm.beta << std::endl;
  if (c.alg == TRACE_APMLR) {
    std::cout << "Beta 1=" << c.prm.beta1 << std::endl;
    std::cout << "Delta 0=" << c.prm.delta << std::endl;
    std::cout << "Delta 1=" << c.prm.delta1 << std::endl;
    std::cout << "Regw=" << c.prm.regw << std::endl;
  }
  return c;
}

}  // namespace

int main(int argc, char **argv) {
  const Config c = kApp == kUnified ? parse_unified(argc, argv) : parse_main(argc, argv);
  if (c.iters < 0) die("the number of iterations must be >= 0");

  // ---- read (sirt_main.cc:96-118: ReadMetadata, ReadSlices, ReadTheta)
  const auto t_read = std::chrono::steady_clock::now();
  int nd = 0;
  int64_t dims[TRACE_H5_MAX_DIMS];
  h5_check(trace_h5_dims(c.proj_file.c_str(), c.proj_dset.c_str(), &nd, dims),
           "projection dataset");
  if (nd != 3) die("the projection dataset must be 3-D [projections][slices][columns]");
  const int P = (int)dims[0], S = (int)dims[1], C = (int)dims[2];
  if (P < 1 || S < 1 || C < 1) die("the projection dataset is empty");
  std::vector<float> sino((size_t)P * S * C);
  h5_check(trace_h5_read_slices(c.proj_file.c_str(), c.proj_dset.c_str(), 0, S, sino.data()),
           "projection dataset");
  int tnd = 0;
  int64_t tdims[TRACE_H5_MAX_DIMS];
  h5_check(trace_h5_dims(c.theta_file.c_str(), c.theta_dset.c_str(), &tnd, tdims),
           "theta dataset");
  if (tnd != 1) die("Theta file dimension should be 1!");  // trace_h5io.cc:89-90
  if (tdims[0] != P)
    die("theta holds " + std::to_string(tdims[0]) + " angles, the projections " +
        std::to_string(P));
  std::vector<float> theta(P);
  h5_check(trace_h5_read(c.theta_file.c_str(), c.theta_dset.c_str(), theta.data(), P),
           "theta dataset");
  const double read_s = seconds_since(t_read);

  // ---- preprocessing (a per-program choice: sirt_main.cc:120-121,
  // mlem_main.cc:113-116, pml_main.cc:121-124, apmlr_main.cc:151-154)
  if (c.deg2rad) trace_deg2rad(theta.data(), theta.data(), P);
  if (c.alg != TRACE_SIRT) trace_absolute(sino.data(), (int64_t)sino.size());
  const float init = c.alg == TRACE_SIRT ? 0.f : 1.f;  // recon_init_val of each main

  // ---- reconstruct
  const auto t_rec = std::chrono::steady_clock::now();
  hip_check(hipSetDevice(c.device), "hipSetDevice");
  trace_geom_t g{P, S, C, C, c.center, theta.data()};
  trace_handle_t *h = nullptr;
  trace_check(trace_create(&h, &g, c.device, nullptr), "trace_create");
  const size_t nrec = (size_t)S * C * C;
  std::vector<float> recon(nrec, init);
  float *d_sino = nullptr, *d_recon = nullptr;
  hip_check(hipMalloc(&d_sino, sino.size() * sizeof(float)), "hipMalloc");
  hip_check(hipMalloc(&d_recon, nrec * sizeof(float)), "hipMalloc");
  hip_check(hipMemcpy(d_sino, sino.data(), sino.size() * sizeof(float), hipMemcpyHostToDevice),
            "copy sinogram");
  hip_check(hipMemcpy(d_recon, recon.data(), nrec * sizeof(float), hipMemcpyHostToDevice),
            "copy initial image");
  for (int i = 0; i < c.iters; ++i) {
    std::cout << "Iteration: " << i << std::endl;  // sirt_main.cc:178
    trace_check(trace_reconstruct(h, c.alg, 1, d_sino, d_recon, &c.prm), "trace_reconstruct");
  }
  hip_check(hipMemcpy(recon.data(), d_recon, nrec * sizeof(float), hipMemcpyDeviceToHost),
            "copy result");
  (void)hipFree(d_sino);
  (void)hipFree(d_recon);
  trace_destroy(h);
  const double rec_s = seconds_since(t_rec);

  // ---- write [S][N][N] (WriteRecon, trace_h5io.cc:470-503)
  const auto t_write = std::chrono::steady_clock::now();
  const int64_t odims[3] = {S, C, C};
  h5_check(trace_h5_write(c.out_file.c_str(), c.out_dset.c_str(), 3, odims, TRACE_H5_F32,
                          recon.data(), 1),
           "output");
  const double write_s = seconds_since(t_write);

  // phase timers (sirt_main.cc:216-222)
  std::cout << "Reconstruction time=" << rec_s << std::endl;
  std::cout << "Read time=" << read_s << std::endl;
  std::cout << "Write time=" << write_s << std::endl;
  return 0;
}
