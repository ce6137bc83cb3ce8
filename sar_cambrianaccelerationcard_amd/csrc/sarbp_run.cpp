// sarbp_run.cpp -- command-line imager with the host flow of the reference's
// SAR_Processor (code/SAR_BP/CUDA/main.cpp, BPA_CUDA.cpp:11-100): build the
// scene, simulate its raw echoes (generateEchoData, :78-81), move them to the
// device, range-compress and back-project them (applyBPAlgorithm, :83-90),
// copy the image back and save it as raw complex64 (saveResult, :92-100) with
// a JSON sidecar.  Where the reference exits on a CUDA error (CHECK_CUDA,
// BPA_CUDA.h:11-17), this prints the library's status text and returns 1.
//
//   sarbp_run [--pulses 512] [--samples 2048] [--nx 512] [--ny 512]
//             [--spacing 0.5] [--interp nearest|linear|sinc8]
//             [--precision ref|f64|f32] [--targets 3] [--target-spacing 20]
//             [--snr-db 20 | --no-noise] [--device 0] [--out sar_image.bin]
//   sarbp_run --reference-scene [--rg0 10000] [--h 5000] [--la 4] [--mslow 1024]
//             [--nfast 300] [--device 0] [--out sar_image.bin]
//
// Scene: X band (9.6 GHz), LFM 150 MHz / 2 us, 200 MHz complex sampling, 4x
// range upsampling; straight track at 150 m/s, 3000 m altitude, 1 kHz PRF,
// centred on the scene; a K x K grid of unit point targets around 4000 m
// ground range on image pixels.
//
// --reference-scene runs the reference's own chain with its own inputs: the
// SARParams defaults (BPA_CUDA.h:28-50, overridable), the five targets and
// the slow/fast-time axes of initializeArrays (BPA_CUDA.cpp:25-48), the echo
// on the device (sarbp_launch_generate_echo for launchGenerateEcho) and the
// back-projection onto the [Mslow][Nfast] image (sarbp_launch_backprojection
// for launchBackProjection, BPA_CUDA.cpp:84).  With the defaults the fast-time
// axis ends 2.5 us after transmission while the targets' echoes arrive after
// ~67 us, so the echo and the image are zero, as the reference's are.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "sarbp.h"
#include "sarsim.h"

namespace {

struct Options {
  int pulses = 512, samples = 2048, nx = 512, ny = 512, targets = 3, device = 0;
  double spacing = 0.5, target_spacing = 20.0, snr_db = 20.0;
  bool noise = true;
  bool reference_scene = false;
  float rg0 = 10000.0f, h = 5000.0f, la = 4.0f;
  int mslow = 1024, nfast = -1;  // nfast -1: (int)(Tr * Fr) as SARParams computes it
  std::string interp = "linear", precision = "ref", out = "sar_image.bin";
};

void usage() {
  std::fprintf(stderr,
               "usage: sarbp_run [--pulses N] [--samples N] [--nx N] [--ny N] [--spacing M]\n"
               "                 [--interp nearest|linear|sinc8] [--precision ref|f64|f32]\n"
               "                 [--targets K] [--target-spacing M] [--snr-db DB | --no-noise]\n"
               "                 [--device D] [--out FILE]\n"
               "       sarbp_run --reference-scene [--rg0 M] [--h M] [--la M] [--mslow N] [--nfast N]\n"
               "                 [--device D] [--out FILE]\n");
}

bool parse(int argc, char** argv, Options* o) {
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&](const char* name) -> const char* {
      if (i + 1 >= argc) throw std::runtime_error(std::string("missing value for ") + name);
      return argv[++i];
    };
    if (a == "--pulses") o->pulses = std::atoi(next("--pulses"));
    else if (a == "--samples") o->samples = std::atoi(next("--samples"));
    else if (a == "--nx") o->nx = std::atoi(next("--nx"));
    else if (a == "--ny") o->ny = std::atoi(next("--ny"));
    else if (a == "--spacing") o->spacing = std::atof(next("--spacing"));
    else if (a == "--interp") o->interp = next("--interp");
    else if (a == "--precision") o->precision = next("--precision");
    else if (a == "--targets") o->targets = std::atoi(next("--targets"));
    else if (a == "--target-spacing") o->target_spacing = std::atof(next("--target-spacing"));
    else if (a == "--snr-db") o->snr_db = std::atof(next("--snr-db"));
    else if (a == "--no-noise") o->noise = false;
    else if (a == "--device") o->device = std::atoi(next("--device"));
    else if (a == "--reference-scene") o->reference_scene = true;
    else if (a == "--rg0") o->rg0 = (float)std::atof(next("--rg0"));
    else if (a == "--h") o->h = (float)std::atof(next("--h"));
    else if (a == "--la") o->la = (float)std::atof(next("--la"));
    else if (a == "--mslow") o->mslow = std::atoi(next("--mslow"));
    else if (a == "--nfast") o->nfast = std::atoi(next("--nfast"));
    else if (a == "--out") o->out = next("--out");
    else if (a == "-h" || a == "--help") return false;
    else throw std::runtime_error("unknown option " + a);
  }
  if (o->pulses < 1 || o->samples < 16 || o->nx < 1 || o->ny < 1 || !(o->spacing > 0) || o->targets < 0)
    throw std::runtime_error("sizes must be positive");
  if (o->mslow < 2 || o->nfast == 0 || o->nfast < -1) throw std::runtime_error("--mslow >= 2 and --nfast >= 1");
  return true;
}

void check(sarbp_status st, const char* what) {
  if (st != SARBP_OK)
    throw std::runtime_error(std::string(what) + ": " + sarbp_strerror(st) + ": " + sarbp_last_error_detail());
}

void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

int enum_of(const std::string& v, const char* const* names, int n, const char* what) {
  for (int i = 0; i < n; ++i)
    if (v == names[i]) return i;
  throw std::runtime_error(std::string("bad ") + what + " '" + v + "'");
}

template <class T>
struct DeviceBuffer {
  T* p = nullptr;
  explicit DeviceBuffer(size_t n) { check_hip(hipMalloc(&p, n * sizeof(T) + 1), "hipMalloc"); }
  ~DeviceBuffer() { (void)hipFree(p); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
};

// The reference's SAR_Processor, on this library.
class SarProcessor {
 public:
  explicit SarProcessor(const Options& o) : o_(o) {
    check(sarbp_create(&ctx_, o.device), "sarbp_create");
    const double c = 299792458.0, fs = 200e6;
    radar_.c = c;
    radar_.fc = 9.6e9;
    radar_.bandwidth = 150e6;
    radar_.pulse_width = 2e-6;
    radar_.fs = fs;
    radar_.upsample = 4;
    radar_.crop_offset = 0;
    radar_.n_out = o.samples;
    radar_.beam_half_width = 0.0;
    // Raw window centred on the scene-centre slant range from the track centre.
    const int span = (o.samples + radar_.upsample - 1) / radar_.upsample;
    const int chirp = (int)std::ceil(radar_.pulse_width * fs);
    radar_.n_raw = span + chirp;
    const double rc = std::hypot(kGroundRange, kAltitude);
    radar_.t0 = 2.0 * (rc - 0.5 * span * (0.5 * c / fs)) / c;
    pos_.assign((size_t)3 * o.pulses, 0.0);
    sarsim_track_straight(pos_.data(), o.pulses, kVelocity, kPrf, 0.0, 0.0, kAltitude);
    const double off = 0.5 * (o.targets - 1);
    for (int a = 0; a < o.targets; ++a)
      for (int b = 0; b < o.targets; ++b) {
        const double t[5] = {(b - off) * o.target_spacing, kGroundRange + (a - off) * o.target_spacing, 0.0, 1.0,
                             0.0};
        targets_.insert(targets_.end(), t, t + 5);
      }
    grid_ = {{-(o.nx / 2) * o.spacing, kGroundRange - (o.ny / 2) * o.spacing, 0.0},
             {o.spacing, 0.0, 0.0},
             {0.0, o.spacing, 0.0},
             o.nx,
             o.ny};
  }
  ~SarProcessor() { sarbp_destroy(ctx_); }

  // generateEchoData + applyBPAlgorithm (BPA_CUDA.cpp:70-90).
  void process() {
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<double> raw64((size_t)2 * o_.pulses * radar_.n_raw, 0.0);
    sarsim_echo_raw(pos_.data(), o_.pulses, targets_.data(), (int32_t)(targets_.size() / 5), &radar_,
                    raw64.data(), 0);
    if (o_.noise) sarsim_add_noise(raw64.data(), (int64_t)o_.pulses * radar_.n_raw, std::pow(10.0, -o_.snr_db / 10),
                                   20260001);
    std::vector<float> raw32(raw64.size());
    for (size_t i = 0; i < raw64.size(); ++i) raw32[i] = (float)raw64[i];
    const auto t1 = std::chrono::steady_clock::now();

    sarbp_range_params rp{};
    rp.c = radar_.c;
    rp.bandwidth = radar_.bandwidth;
    rp.pulse_width = radar_.pulse_width;
    rp.fs = radar_.fs;
    rp.t0 = radar_.t0;
    rp.n_raw = radar_.n_raw;
    rp.upsample = radar_.upsample;
    rp.crop_offset = 0;
    rp.n_out = o_.samples;
    rp.dtype = SARBP_DTYPE_C64;
    double r0 = 0, dr = 0;
    check(sarbp_range_info(&rp, nullptr, &r0, &dr), "sarbp_range_info");

    static const char* const kInterp[] = {"nearest", "linear", "sinc8"};
    static const char* const kPrec[] = {"ref", "f64", "f32"};
    sarbp_params p{};
    p.c = radar_.c;
    p.fc = radar_.fc;
    p.r0 = r0;
    p.dr = dr;
    p.kaiser_beta = 6.0;
    p.n_pulses = o_.pulses;
    p.n_samples = o_.samples;
    p.interp = enum_of(o_.interp, kInterp, 3, "interp");
    p.precision = enum_of(o_.precision, kPrec, 3, "precision");
    p.dtype = SARBP_DTYPE_C64;
    p.accumulate = 0;

    const size_t npix = (size_t)o_.nx * o_.ny;
    DeviceBuffer<sarbp_cf32> d_raw((size_t)o_.pulses * radar_.n_raw), d_prof((size_t)o_.pulses * o_.samples),
        d_img(npix);
    DeviceBuffer<double> d_pos(pos_.size());
    check_hip(hipMemcpy(d_raw.p, raw32.data(), raw32.size() * sizeof(float), hipMemcpyHostToDevice), "H2D raw");
    check_hip(hipMemcpy(d_pos.p, pos_.data(), pos_.size() * sizeof(double), hipMemcpyHostToDevice), "H2D pos");
    const auto t2 = std::chrono::steady_clock::now();
    check(sarbp_range_compress(ctx_, d_raw.p, o_.pulses, &rp, d_prof.p, nullptr), "sarbp_range_compress");
    check(sarbp_backproject_sync(ctx_, d_prof.p, d_pos.p, &grid_, &p, d_img.p, nullptr), "sarbp_backproject");
    const auto t3 = std::chrono::steady_clock::now();
    image_.resize(npix);
    check_hip(hipMemcpy(image_.data(), d_img.p, npix * sizeof(sarbp_cf32), hipMemcpyDeviceToHost), "D2H image");
    const auto t4 = std::chrono::steady_clock::now();
    auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
    std::printf("echoes %.1f ms, H2D %.1f ms, range compression + BP %.1f ms, D2H %.1f ms: "
                "%.3e pixel-pulses/s on the device\n",
                ms(t0, t1), ms(t1, t2), ms(t2, t3), ms(t3, t4), (double)npix * o_.pulses / (ms(t2, t3) * 1e-3));
    r0_ = r0;
    dr_ = dr;
  }

  // saveResult (BPA_CUDA.cpp:92-100) + a JSON sidecar.
  void save(const std::string& path) const {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot open " + path);
    const size_t n = std::fwrite(image_.data(), sizeof(sarbp_cf32), image_.size(), f);
    std::fclose(f);
    if (n != image_.size()) throw std::runtime_error("short write to " + path);
    FILE* j = std::fopen((path + ".json").c_str(), "w");
    if (!j) throw std::runtime_error("cannot open " + path + ".json");
    std::fprintf(j,
                 "{\"format\": \"raw complex64 interleaved (re, im), row-major [ny][nx]\", \"ny\": %d, "
                 "\"nx\": %d, \"dtype\": \"complex64\", \"bytes\": %zu, \"meta\": {\"writer\": \"sarbp_run\", "
                 "\"pulses\": %d, \"samples\": %d, \"spacing_m\": %.6g, \"origin\": [%.6f, %.6f, %.6f], "
                 "\"interp\": \"%s\", \"precision\": \"%s\", \"r0_m\": %.9f, \"dr_m\": %.12g, \"targets\": %d, "
                 "\"target_spacing_m\": %.6g, \"noise\": %s}}\n",
                 o_.ny, o_.nx, image_.size() * sizeof(sarbp_cf32), o_.pulses, o_.samples, o_.spacing,
                 grid_.origin[0], grid_.origin[1], grid_.origin[2], o_.interp.c_str(), o_.precision.c_str(), r0_,
                 dr_, o_.targets, o_.target_spacing, o_.noise ? "true" : "false");
    std::fclose(j);
  }

  // Peak magnitude at each target's pixel (a unit target focuses to
  // pulses x Tp fs there).
  void report() const {
    const double expect = (double)o_.pulses * radar_.pulse_width * radar_.fs;
    for (size_t k = 0; k < targets_.size() / 5; ++k) {
      const int ix = (int)std::lround((targets_[5 * k] - grid_.origin[0]) / o_.spacing);
      const int iy = (int)std::lround((targets_[5 * k + 1] - grid_.origin[1]) / o_.spacing);
      if (ix < 0 || iy < 0 || ix >= o_.nx || iy >= o_.ny) continue;
      const sarbp_cf32 v = image_[(size_t)iy * o_.nx + ix];
      std::printf("target %zu at pixel (%d, %d): |I| = %.1f (%.3f of pulses x Tp fs)\n", k, ix, iy,
                  std::hypot(v.re, v.im), std::hypot(v.re, v.im) / expect);
    }
  }

 private:
  static constexpr double kVelocity = 150.0, kPrf = 1000.0, kAltitude = 3000.0, kGroundRange = 4000.0;
  Options o_;
  sarbp_context* ctx_ = nullptr;
  sarsim_radar radar_{};
  std::vector<double> pos_, targets_;
  sarbp_grid grid_{};
  std::vector<sarbp_cf32> image_;
  double r0_ = 0, dr_ = 0;
};

// The reference's SAR_Processor with its own inputs (BPA_CUDA.cpp:11-100): the
// axes in float exactly as initializeArrays computes them.
int run_reference_scene(const Options& o) {
  check_hip(hipSetDevice(o.device), "hipSetDevice");
  sarbp_sar_params p{};
  p.c = 3e8f;
  p.fc = 5.3e9f;
  p.Kr = 4e12f;
  p.Tr = 2.5e-6f;
  p.Br = 1e8f;
  p.Fr = 1.2e8f;
  p.vr = 150.0f;
  p.H = o.h;
  p.La = o.la;
  p.Rg0 = o.rg0;
  p.Az0 = 0.0f;
  p.lambda = p.c / p.fc;
  p.Nfast = o.nfast > 0 ? o.nfast : static_cast<int>(p.Tr * p.Fr);
  p.Mslow = o.mslow;
  p.Ntarget = 5;
  const float pt[15] = {p.Az0 - 10, p.Rg0 - 20, 1.0f, p.Az0 + 20, p.Rg0 + 30, 0.8f, p.Az0 - 30, p.Rg0 + 10, 1.2f,
                        p.Az0 + 40, p.Rg0 - 40, 0.9f, p.Az0,      p.Rg0,      1.5f};
  const float ta_start = -p.La / (2 * p.vr), ta_end = p.La / (2 * p.vr);
  const float ta_step = (ta_end - ta_start) / (p.Mslow - 1);
  std::vector<float> ta((size_t)p.Mslow), tr((size_t)p.Nfast);
  for (int i = 0; i < p.Mslow; ++i) ta[i] = ta_start + i * ta_step;
  const float tr_step = 1.0f / p.Fr;
  for (int i = 0; i < p.Nfast; ++i) tr[i] = i * tr_step;
  const size_t n = (size_t)p.Mslow * p.Nfast;
  DeviceBuffer<sarbp_cf32> d_srnm(n), d_img(n);
  DeviceBuffer<float> d_pt(15), d_ta(ta.size()), d_tr(tr.size());
  check_hip(hipMemcpy(d_pt.p, pt, sizeof pt, hipMemcpyHostToDevice), "H2D targets");
  check_hip(hipMemcpy(d_ta.p, ta.data(), ta.size() * sizeof(float), hipMemcpyHostToDevice), "H2D ta");
  check_hip(hipMemcpy(d_tr.p, tr.data(), tr.size() * sizeof(float), hipMemcpyHostToDevice), "H2D tr");
  const auto t0 = std::chrono::steady_clock::now();
  check(sarbp_launch_generate_echo(d_srnm.p, d_pt.p, d_ta.p, d_tr.p, p, p.Mslow, p.Nfast),
        "sarbp_launch_generate_echo");
  const auto t1 = std::chrono::steady_clock::now();
  check(sarbp_launch_backprojection(d_srnm.p, d_img.p, p, p.Mslow, p.Nfast), "sarbp_launch_backprojection");
  const auto t2 = std::chrono::steady_clock::now();
  std::vector<sarbp_cf32> echo(n), img(n);
  check_hip(hipMemcpy(echo.data(), d_srnm.p, n * sizeof(sarbp_cf32), hipMemcpyDeviceToHost), "D2H echo");
  check_hip(hipMemcpy(img.data(), d_img.p, n * sizeof(sarbp_cf32), hipMemcpyDeviceToHost), "D2H image");
  size_t nz_echo = 0;
  double peak = 0.0;
  for (size_t i = 0; i < n; ++i) {
    nz_echo += (echo[i].re != 0.0f || echo[i].im != 0.0f);
    peak = std::max(peak, std::hypot((double)img[i].re, (double)img[i].im));
  }
  auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
  std::printf("reference scene: %d x %d, echo %.2f ms (%zu non-zero samples), BP %.2f ms, image peak |I| = %.4g\n",
              p.Mslow, p.Nfast, ms(t0, t1), nz_echo, ms(t1, t2), peak);
  FILE* f = std::fopen(o.out.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot open " + o.out);
  const size_t w = std::fwrite(img.data(), sizeof(sarbp_cf32), n, f);
  std::fclose(f);
  if (w != n) throw std::runtime_error("short write to " + o.out);
  FILE* j = std::fopen((o.out + ".json").c_str(), "w");
  if (!j) throw std::runtime_error("cannot open " + o.out + ".json");
  std::fprintf(j,
               "{\"format\": \"raw complex64 interleaved (re, im), row-major [ny][nx]\", \"ny\": %d, \"nx\": %d, "
               "\"dtype\": \"complex64\", \"bytes\": %zu, \"meta\": {\"writer\": \"sarbp_run --reference-scene\", "
               "\"Rg0\": %.6g, \"H\": %.6g, \"La\": %.6g, \"Mslow\": %d, \"Nfast\": %d, \"echo_nonzero\": %zu}}\n",
               p.Mslow, p.Nfast, n * sizeof(sarbp_cf32), p.Rg0, p.H, p.La, p.Mslow, p.Nfast, nz_echo);
  std::fclose(j);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    Options o;
    if (!parse(argc, argv, &o)) {
      usage();
      return 2;
    }
    if (o.reference_scene) return run_reference_scene(o);
    SarProcessor processor(o);
    processor.process();
    processor.save(o.out);
    processor.report();
    std::printf("SAR processing completed: %s\n", o.out.c_str());
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "Error: %s\n", e.what());
    return 1;
  }
}
