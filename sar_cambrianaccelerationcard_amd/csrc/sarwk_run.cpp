// sarwk_run.cpp -- command-line omega-K imager with the host flow of the
// reference's process_wk (code/SAR_WK/wk_v5.2.3/wk.cc:43-343, main :345-358):
// read the run parameters ("name: value" lines, readParams read.cc:54-82) and
// the echo (whitespace-separated "re im" pairs, readComplexData read.cc:15-32;
// or raw complex64 for a .bin name), check its size (wk.cc:81-84), image it on
// the device through sarbp_omega_k (padding and down-conversion, 2-D FFT,
// reference function, sinc Stolt, 2-D inverse FFT), print the first 5 x 5
// values (wk.cc:295-303) and, with --out, write the image: the result.txt
// text format of wk.cc:304-332 for a .txt name, else raw complex64 with a
// JSON sidecar.  Where the reference throws on a CNNL/CNRT error and logs it
// (wk.cc:338-341), this prints the library's status text and returns 1.
//
//   sarwk_run [--params data/paras.txt] [--echo data/echo.txt] [--na 3072] [--nr 2048]
//             [--vr 7062] [--f-nc -6900] [--taps 6] [--window floor|centered]
//             [--range-ref r0|centre] [--device 0] [--info | -i] [--out | -o [FILE]]
//
// Vr and f_nc default to the constants of wk.cc:62-64; Kr is taken
// as stored in the parameter file (signed: negative for a down-chirp).
// --range-ref centre (not in the reference) takes R0 as the slant range of
// the first echo sample and references the window's centre, so targets in
// the far half of the range window focus too.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "sarbp.h"

namespace {

struct Options {
  std::string params = "data/paras.txt", echo = "data/echo.txt", out;
  int na = 3072, nr = 2048, taps = 6, device = 0;
  double vr = 7062.0, f_nc = -6900.0;
  std::string window = "floor", range_ref = "r0";
  bool info = false, write = false;
};

void usage() {
  std::fprintf(stderr,
               "usage: sarwk_run [--params FILE] [--echo FILE] [--na N] [--nr N] [--vr V] [--f-nc F]\n"
               "                 [--taps P] [--window floor|centered] [--range-ref r0|centre]\n"
               "                 [--device D] [--info|-i]\n"
               "                 [--out|-o [FILE]]\n");
}

bool parse(int argc, char** argv, Options* o) {
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&](const char* name) -> const char* {
      if (i + 1 >= argc) throw std::runtime_error(std::string("missing value for ") + name);
      return argv[++i];
    };
    if (a == "--params") o->params = next("--params");
    else if (a == "--echo") o->echo = next("--echo");
    else if (a == "--na") o->na = std::atoi(next("--na"));
    else if (a == "--nr") o->nr = std::atoi(next("--nr"));
    else if (a == "--vr") o->vr = std::atof(next("--vr"));
    else if (a == "--f-nc") o->f_nc = std::atof(next("--f-nc"));
    else if (a == "--taps") o->taps = std::atoi(next("--taps"));
    else if (a == "--window") o->window = next("--window");
    else if (a == "--range-ref") o->range_ref = next("--range-ref");
    else if (a == "--device") o->device = std::atoi(next("--device"));
    else if (a == "--info" || a == "-i") o->info = true;
    else if (a == "--out" || a == "-o") {
      o->write = true;
      if (i + 1 < argc && argv[i + 1][0] != '-') o->out = argv[++i];
    } else if (a == "-h" || a == "--help") return false;
    else throw std::runtime_error("unknown option " + a);
  }
  if (o->out.empty()) o->out = "result.txt";
  if (o->na < 1 || o->nr < 1) throw std::runtime_error("--na and --nr must be positive");
  if (o->window != "floor" && o->window != "centered") throw std::runtime_error("--window floor|centered");
  if (o->range_ref != "r0" && o->range_ref != "centre") throw std::runtime_error("--range-ref r0|centre");
  return true;
}

void check(sarbp_status st, const char* what) {
  if (st != SARBP_OK)
    throw std::runtime_error(std::string(what) + ": " + sarbp_strerror(st) + ": " + sarbp_last_error_detail());
}

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

bool ends_with(const std::string& s, const char* suffix) {
  const size_t n = std::strlen(suffix);
  return s.size() >= n && s.compare(s.size() - n, n, suffix) == 0;
}

// "name: value" lines; unknown names and lines without ':' are skipped.
std::map<std::string, double> read_params(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open parameter file " + path);
  std::map<std::string, double> m;
  std::string line;
  while (std::getline(f, line)) {
    const size_t p = line.find(':');
    if (p == std::string::npos) continue;
    m[line.substr(0, p)] = std::stod(line.substr(p + 1));
  }
  for (const char* k : {"Fr", "PRF", "f0", "R0", "Kr", "c"})
    if (!m.count(k)) throw std::runtime_error(std::string("parameter ") + k + " missing in " + path);
  return m;
}

std::vector<sarbp_cf32> read_echo(const std::string& path) {
  std::vector<sarbp_cf32> v;
  if (ends_with(path, ".bin")) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open echo file " + path);
    std::fseek(f, 0, SEEK_END);
    const long bytes = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    v.resize((size_t)bytes / sizeof(sarbp_cf32));
    const size_t got = std::fread(v.data(), sizeof(sarbp_cf32), v.size(), f);
    std::fclose(f);
    if (got != v.size()) throw std::runtime_error("short read from " + path);
    return v;
  }
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open echo file " + path);
  float re, im;
  while (f >> re >> im) v.push_back({re, im});
  return v;
}

void write_image(const std::string& path, const std::vector<sarbp_cf32>& img, int na, int nr,
                 const sarbp_wk_params& wp) {
  if (ends_with(path, ".txt")) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("cannot open " + path);
    for (int i = 0; i < na; ++i) {
      for (int j = 0; j < nr; ++j) {
        const sarbp_cf32 v = img[(size_t)i * nr + j];
        std::fprintf(f, "%.6e %.6e%s", v.re, v.im, j < nr - 1 ? " " : "");
      }
      std::fputc('\n', f);
    }
    std::fclose(f);
    return;
  }
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot open " + path);
  const size_t n = std::fwrite(img.data(), sizeof(sarbp_cf32), img.size(), f);
  std::fclose(f);
  if (n != img.size()) throw std::runtime_error("short write to " + path);
  FILE* j = std::fopen((path + ".json").c_str(), "w");
  if (!j) throw std::runtime_error("cannot open " + path + ".json");
  std::fprintf(j,
               "{\"format\": \"raw complex64 interleaved (re, im), row-major [ny][nx]\", \"ny\": %d, "
               "\"nx\": %d, \"dtype\": \"complex64\", \"bytes\": %zu, \"meta\": {\"writer\": \"sarwk_run\", "
               "\"rows\": \"azimuth frequency lines after the inverse transform\", \"c\": %.10g, \"f0\": %.10g, "
               "\"fs\": %.10g, \"prf\": %.10g, \"kr\": %.10g, \"vr\": %.10g, \"r0\": %.10g, \"f_nc\": %.10g, "
               "\"taps\": %d, \"window\": \"%s\", \"range_ref\": \"%s\"}}\n",
               na, nr, img.size() * sizeof(sarbp_cf32), wp.c, wp.f0, wp.fs, wp.prf, wp.kr, wp.vr, wp.r0, wp.f_nc,
               wp.taps, wp.window == SARBP_STOLT_FLOOR ? "floor" : "centered",
               wp.range_ref == SARBP_WK_RANGE_REF_CENTRE ? "centre" : "r0");
  std::fclose(j);
}

}  // namespace

int main(int argc, char** argv) {
  sarbp_context* ctx = nullptr;
  void *d_echo = nullptr, *d_img = nullptr;
  int rc = 0;
  try {
    Options o;
    if (!parse(argc, argv, &o)) {
      usage();
      return 2;
    }
    const auto prm = read_params(o.params);
    sarbp_wk_params wp{};
    wp.c = prm.at("c");
    wp.f0 = prm.at("f0");
    wp.fs = prm.at("Fr");
    wp.prf = prm.at("PRF");
    wp.kr = prm.at("Kr");
    wp.r0 = prm.at("R0");
    wp.vr = o.vr;
    wp.f_nc = o.f_nc;
    wp.taps = o.taps;
    wp.window = o.window == "floor" ? SARBP_STOLT_FLOOR : SARBP_STOLT_CENTERED;
    wp.range_ref = o.range_ref == "centre" ? SARBP_WK_RANGE_REF_CENTRE : SARBP_WK_RANGE_REF_R0;
    if (o.info)
      std::printf("Fr %.6g Hz, PRF %.6g Hz, f0 %.6g Hz, R0 %.6f m, Kr %.6g Hz/s, c %.6g m/s, Vr %.6g m/s, "
                  "f_nc %.6g Hz, P %d, window %s, range reference %s\n",
                  wp.fs, wp.prf, wp.f0, wp.r0, wp.kr, wp.c, wp.vr, wp.f_nc, wp.taps, o.window.c_str(),
                  o.range_ref.c_str());
    const auto echo = read_echo(o.echo);
    if (echo.size() != (size_t)o.na * o.nr)
      throw std::runtime_error("echo holds " + std::to_string(echo.size()) + " samples, expected " +
                               std::to_string((size_t)o.na * o.nr));
    int32_t na_p = 0, nr_p = 0;
    check(sarbp_wk_padded_size(o.na, o.nr, &na_p, &nr_p), "sarbp_wk_padded_size");
    std::printf("echo %d x %d, padded %d x %d\n", o.na, o.nr, na_p, nr_p);

    hip_check(hipSetDevice(o.device), "hipSetDevice");
    check(sarbp_create(&ctx, o.device), "sarbp_create");
    const size_t in_bytes = echo.size() * sizeof(sarbp_cf32);
    const size_t out_n = (size_t)na_p * nr_p;
    hip_check(hipMalloc(&d_echo, in_bytes), "hipMalloc(echo)");
    hip_check(hipMalloc(&d_img, out_n * sizeof(sarbp_cf32)), "hipMalloc(image)");
    const auto t0 = std::chrono::steady_clock::now();
    hip_check(hipMemcpy(d_echo, echo.data(), in_bytes, hipMemcpyHostToDevice), "H2D");
    const auto t1 = std::chrono::steady_clock::now();
    // The first call of a size builds its transform plan; the second is timed.
    for (int pass = 0; pass < 2; ++pass) {
      check(sarbp_omega_k(ctx, static_cast<const sarbp_cf32*>(d_echo), o.na, o.nr, &wp,
                          static_cast<sarbp_cf32*>(d_img), nullptr),
            "sarbp_omega_k");
      hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    }
    const auto t2 = std::chrono::steady_clock::now();
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    hip_check(hipEventRecord(e0, nullptr), "hipEventRecord");
    check(sarbp_omega_k(ctx, static_cast<const sarbp_cf32*>(d_echo), o.na, o.nr, &wp,
                        static_cast<sarbp_cf32*>(d_img), nullptr),
          "sarbp_omega_k");
    hip_check(hipEventRecord(e1, nullptr), "hipEventRecord");
    hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
    float kms = 0.0f;
    hip_check(hipEventElapsedTime(&kms, e0, e1), "hipEventElapsedTime");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::vector<sarbp_cf32> img(out_n);
    hip_check(hipMemcpy(img.data(), d_img, out_n * sizeof(sarbp_cf32), hipMemcpyDeviceToHost), "D2H");
    const auto t3 = std::chrono::steady_clock::now();
    auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
    std::printf("H2D %.2f ms, first two calls %.1f ms, omega-K %.3f ms on the device (%.3e echo samples/s)\n",
                ms(t0, t1), ms(t1, t2), kms, (double)o.na * o.nr / (kms * 1e-3));
    std::printf("first 5 x 5 values:\n");
    for (int i = 0; i < 5 && i < na_p; ++i) {
      for (int j = 0; j < 5 && j < nr_p; ++j) {
        const sarbp_cf32 v = img[(size_t)i * nr_p + j];
        std::printf("(%.3e,%.3e) ", v.re, v.im);
      }
      std::printf("\n");
    }
    if (o.write) {
      write_image(o.out, img, na_p, nr_p, wp);
      std::printf("image written to %s (total %.1f ms)\n", o.out.c_str(), ms(t0, std::chrono::steady_clock::now()));
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "sarwk_run: %s\n", e.what());
    rc = 1;
  }
  if (d_echo) (void)hipFree(d_echo);
  if (d_img) (void)hipFree(d_img);
  sarbp_destroy(ctx);
  return rc;
}
