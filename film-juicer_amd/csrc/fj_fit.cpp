// fj_fit.cpp — host side of the differentiable print route: the record of a
// parameter vector, the derivative tables of the host preparation, and the
// forward / backward launch sequences on the caller's stream.
#include "fj_fit.h"

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fj/fj_mi355x_ext.h"
#include "fj_fft.h"
#include "fj_host.h"
#include "fj_owner.h"
#include "fj_plan.h"
#include "fj_spatial.h"

namespace fjx {

void dichroic_filters(float filters[81][3]);  // fj_print_prep.cpp

namespace {

// The record of one parameter vector and what the adjoints need besides it.
struct FitRecord {
    FjPreparedHostData rec{};
    std::vector<float> illum;    // 81, filtered main illuminant
    std::vector<float> axes[3];  // compensated axes R, G, B (the record's order)
    std::vector<float> alpha;    // {aM, aY, 0, 0} per compacted print wavelength
    std::vector<float> daxis;    // d axes / d dir_k, [B | G | R]
    float A[9] = {};
    float k_unit[3] = {};        // record donor K, B/G/R
};

bool fit_supported(const FjPreparedHostData& p, int width, int height, std::string& why) {
    const FjFilmExposure& ex = p.film_exposure;
    auto no = [&](const char* what) {
        why = std::string("UnsupportedInput fit ") + what;
        return false;
    };
    if (p.route != FJ_ROUTE_NEGATIVE_PRINT) return no("route other than negative print");
    if (ex.method != FJ_RAW_MALLETT_2019) return no("reconstruction other than Mallett 2019");
    if (ex.flags & (FJ_FILM_DECODE_CCTF | FJ_FILM_AUTO_EXPOSURE)) return no("input decoding or auto exposure");
    if (p.optics.flags != 0) return no("camera or enlarger optics");
    if (p.grain.flags & FJ_GRAIN_ACTIVE) return no("active grain (the increment replaces it)");
    if (p.effects.hash != 0) return no("film effects");
    if (scanner_post_active(p.scanner_post)) return no("scanner post effects or glare");
    if (p.output_color.flags & FJ_COLOR_ENCODE_CCTF) return no("encoded output");
    if (p.dir_couplers.mode != FJ_DIR_NEGATIVE_DONOR_LANGMUIR || p.spatial_dir.hash == 0)
        return no("without negative-donor spatial DIR");
    const FjSpatialDir& sd = p.spatial_dir;
    auto whole = [&](const FjExtent& e) { return e.x == 0 && e.y == 0 && e.width == width && e.height == height; };
    if (sd.recipe_hash != p.dir_couplers.hash || !whole(sd.render_extent) || !whole(sd.full_frame_extent) ||
        !whole(sd.filter_domain_extent)) {
        why = "MalformedPreparedHostData field=spatial_dir.frame";
        return false;
    }
    if (p.film_development.log_exposure.count < 2 || p.print.log_exposure.count < 2)
        return no("curves of fewer than two samples");
    return true;
}

bool finite_params(const FjFitParams& t) {
    return std::isfinite(t.exposure_ev) && std::isfinite(t.print_exposure) && t.print_exposure > 0.0f &&
           std::isfinite(t.filter_y_cc) && std::isfinite(t.filter_m_cc) && std::isfinite(t.film_gamma) &&
           t.film_gamma > 0.0f && std::isfinite(t.print_gamma) && t.print_gamma > 0.0f && std::isfinite(t.dir_k) &&
           t.dir_k > 0.0f;
}

// The dichroic filter factor 1 - (1 - f)(1 - 10^(-cc/100)) (print.rs:636-646)
// and its logarithmic derivative in cc.
void filter_factor(double f, double cc, double& F, double& dlnF) {
    const double t = std::pow(10.0, -cc / 100.0);
    F = 1.0 - (1.0 - f) * (1.0 - t);
    dlnF = (1.0 - f) * (-t * std::log(10.0) / 100.0) / F;
}

// Langmuir v (k + r) / (k + v) in the reference's FP64 order
// (film/dir.rs:109-113) and d/dk.
double langmuir64(double v, double k, double r) { return v * (k + r) / (k + v); }
double langmuir64_dk(double v, double k, double r) { return v * (v - r) / ((k + v) * (k + v)); }

bool build_fit_record(const FjPreparedHostData& nominal, const FjFitParams& th, FitRecord& out, std::string& diag) {
    out.rec = nominal;
    FjPreparedHostData& r = out.rec;
    r.film_exposure.manual_exposure_ev = th.exposure_ev;
    for (float& g : r.film_development.gamma_rgb) g = th.film_gamma;
    r.print.exposure = th.print_exposure;

    // Filtered illuminant: the record's times the filter factors' ratio.
    const FjPrint& pr = nominal.print;
    if (pr.main_illuminant.count != 81 || pr.film_density_cmy.count != 243 || pr.sensitivity_cmy.count != 243 ||
        pr.film_base_density.count != 81) {
        diag = "MalformedPreparedHostData field=print";
        return false;
    }
    float filters[81][3];
    dichroic_filters(filters);
    const double cc0_m = pr.main_cc_cmy[1], cc0_y = pr.main_cc_cmy[2];
    out.illum.resize(81);
    std::vector<double> a_m(81), a_y(81);
    for (int i = 0; i < 81; ++i) {
        double fm0, fm1, fy0, fy1, dm, dy, unused;
        filter_factor(filters[i][1], cc0_m, fm0, unused);
        filter_factor(filters[i][1], cc0_m + th.filter_m_cc, fm1, dm);
        filter_factor(filters[i][2], cc0_y, fy0, unused);
        filter_factor(filters[i][2], cc0_y + th.filter_y_cc, fy1, dy);
        const double ratio = (fm1 * fy1) / (fm0 * fy0);
        out.illum[i] = static_cast<float>(static_cast<double>(pr.main_illuminant.data[i]) * ratio);
        a_m[i] = dm;
        a_y[i] = dy;
    }
    r.print.main_illuminant = FjFloatSpan{out.illum.data(), 81};
    // The compaction of the print spectra (the same wavelengths as the
    // forward's table, fj_plan.cpp pack_print_spectra).
    for (int i = 0; i < 81; ++i) {
        const float* e = pr.film_density_cmy.data + i * 3;
        const float* s = pr.sensitivity_cmy.data + i * 3;
        const float base = pr.film_base_density.data[i];
        const bool any_sens = !std::isnan(s[0]) || !std::isnan(s[1]) || !std::isnan(s[2]);
        if (!any_sens || std::isnan(e[0]) || std::isnan(e[1]) || std::isnan(e[2]) || std::isnan(base) ||
            std::isnan(out.illum[i]))
            continue;
        out.alpha.insert(out.alpha.end(), {static_cast<float>(a_m[i]), static_cast<float>(a_y[i]), 0.0f, 0.0f});
    }
    if (out.alpha.empty()) out.alpha.resize(4, 0.0f);

    // DIR donor K and the compensated axes (film/dir.rs:234-310).
    FjDirCouplers& dc = r.dir_couplers;
    const FjFilmDevelopment& dv = nominal.film_development;
    const size_t n = dv.log_exposure.count;
    if (dv.density_rgb.count != 3 * n || nominal.dir_couplers.compensated_axes_rgb[0].count != n) {
        diag = "MalformedPreparedHostData field=dir_couplers.compensated_axes_rgb";
        return false;
    }
    double k[3], ref[3], unit[3];
    for (int d = 0; d < 3; ++d) {
        unit[d] = nominal.dir_couplers.donor_k_rgb[d];
        dc.donor_k_rgb[d] = th.dir_k == 1.0f ? nominal.dir_couplers.donor_k_rgb[d]
                                             : th.dir_k * nominal.dir_couplers.donor_k_rgb[d];
        k[d] = dc.donor_k_rgb[d];
        ref[d] = dc.density_ref_rgb[d];
        if (!std::isfinite(dc.donor_k_rgb[d]) || !(dc.donor_k_rgb[d] > 0.0f)) {
            diag = "UnsupportedInput fit donor K";
            return false;
        }
    }
    out.daxis.assign(3 * n, 0.0f);
    for (int ch = 0; ch < 3; ++ch) out.axes[ch].resize(n);
    for (size_t s = 0; s < n; ++s) {
        double donors[3], ddk[3];
        for (int d = 0; d < 3; ++d) {
            const double density = dv.density_rgb.data[s * 3 + d];
            donors[d] = langmuir64(density, k[d], ref[d]);
            ddk[d] = langmuir64_dk(density, k[d], ref[d]) * unit[d];
            if (!std::isfinite(donors[d]) || k[d] + density == 0.0) {
                diag = "UnsupportedInput fit DIR donor Langmuir";
                return false;
            }
        }
        for (int rc = 0; rc < 3; ++rc) {
            double correction = 0.0, dcorr = 0.0;
            for (int d = 0; d < 3; ++d) {
                correction += donors[d] * static_cast<double>(dc.matrix_rgb[d * 3 + rc]);
                dcorr += ddk[d] * static_cast<double>(dc.matrix_rgb[d * 3 + rc]);
            }
            const double coordinate = static_cast<double>(dv.log_exposure.data[s]) - correction;
            if (!std::isfinite(coordinate)) {
                diag = "UnsupportedInput fit compensated axis";
                return false;
            }
            out.axes[rc][s] = static_cast<float>(coordinate);
            // packed [B | G | R]: receiver R at 2, G at 1, B at 0 -> slot 2 - rc
            out.daxis[(2 - rc) * n + s] = static_cast<float>(-dcorr);
        }
    }
    for (int ch = 0; ch < 3; ++ch) dc.compensated_axes_rgb[ch] = FjFloatSpan{out.axes[ch].data(), n};
    for (int b = 0; b < 3; ++b) out.k_unit[b] = static_cast<float>(unit[2 - b]);

    // Mallett exposure matrix from the forward's packed table entries
    // (fj_plan.cpp pack_mallett): A[b][j] = sum_i w_ij sens_ib.
    const FjFilmExposure& ex = nominal.film_exposure;
    if (ex.mallett_basis.count != 243 || ex.mallett_illuminant.count != 81 || ex.sensitivity_rgb.count != 243) {
        diag = "MalformedPreparedHostData field=film_exposure.mallett";
        return false;
    }
    double A[9] = {};
    for (int i = 0; i < 81; ++i) {
        const float* bs = ex.mallett_basis.data + i * 3;
        const float ill = ex.mallett_illuminant.data[i];
        const float* sn = ex.sensitivity_rgb.data + i * 3;
        if (!std::isfinite(bs[0]) || !std::isfinite(bs[1]) || !std::isfinite(bs[2]) || !std::isfinite(ill)) continue;
        for (int b = 0; b < 3; ++b) {
            const float sens = std::isfinite(sn[2 - b]) ? sn[2 - b] : 0.0f;
            for (int j = 0; j < 3; ++j)
                A[b * 3 + j] += static_cast<double>(static_cast<float>(double(bs[j]) * ill)) * sens;
        }
    }
    for (int i = 0; i < 9; ++i) out.A[i] = static_cast<float>(A[i]);
    return true;
}

FjStatus refuse(FjErrorBuffer* error, uint32_t category, const std::string& text) {
    return fail(error, category, text);
}

}  // namespace

FjStatus fit_precheck(const FjPreparedHostData& nominal, const FjFitParams& th, int width, int height,
                      FjErrorBuffer* error) {
    uint32_t category = FJ_STATUS_UNSUPPORTED_INPUT;
    std::string text;
    if (width <= 0 || height <= 0) return refuse(error, FJ_STATUS_UNSUPPORTED_INPUT, "UnsupportedInput fit frame extent");
    if (!record_valid(nominal, category, text)) return refuse(error, category, text);
    if (!fit_supported(nominal, width, height, text)) return refuse(error, FJ_STATUS_UNSUPPORTED_INPUT, text);
    if (!finite_params(th) || th.reserved != 0.0f)
        return refuse(error, FJ_STATUS_UNSUPPORTED_INPUT, "UnsupportedInput fit parameters");
    return make_status(FJ_STATUS_SUCCESS);
}

namespace {

// One slot of the device's per-parameter upload ring, held from the
// staging copy to the last launch that reads it; the destructor records the
// slot's event on the stream (after whatever was queued, also on an early
// return), so the slot is reused only after those launches.
struct RingHold {
    DeviceState* dev = nullptr;
    hipStream_t stream = nullptr;
    int index = -1;
    hipError_t acquire(DeviceState& d, hipStream_t s, float*& host, float*& device) {
        dev = &d;
        stream = s;
        if (!d.fit_ring) {
            const size_t bytes = DeviceState::kFitRing * DeviceState::kFitSlotFloats * sizeof(float);
            hipError_t e = hipMalloc(reinterpret_cast<void**>(&d.fit_ring), bytes);
            if (e != hipSuccess) return e;
            e = hipHostMalloc(reinterpret_cast<void**>(&d.fit_staging), bytes, hipHostMallocDefault);
            if (e != hipSuccess) return e;
            for (hipEvent_t& ev : d.fit_events) {
                e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
                if (e != hipSuccess) return e;
            }
        }
        const int i = d.fit_next;
        if (d.fit_pending[i]) {
            const hipError_t e = hipEventSynchronize(d.fit_events[i]);
            if (e != hipSuccess) return e;
            d.fit_pending[i] = false;
        }
        d.fit_next = (i + 1) % DeviceState::kFitRing;
        index = i;
        host = d.fit_staging + i * DeviceState::kFitSlotFloats;
        device = d.fit_ring + i * DeviceState::kFitSlotFloats;
        return hipSuccess;
    }
    ~RingHold() {
        if (index < 0) return;
        if (hipEventRecord(dev->fit_events[index], stream) == hipSuccess) {
            dev->fit_pending[index] = true;
        } else {
            // No event: wait for the stream so the slot is free again.
            (void)hipGetLastError();
            (void)hipStreamSynchronize(stream);
        }
    }
};

size_t aligned64(size_t floats) { return (floats + 63) / 64 * 64; }

// Validation and assembly shared by the forward and the backward.
struct FitSetup {
    FitRecord fr;
    FitArgs args{};
    DctFilterJob job;
    long long total = 0;
    RingHold ring;
};

FjStatus fit_setup(DeviceState& dev, const FjPreparedHostData& nominal, const FjFitParams& th, int width, int height,
                   hipStream_t stream, FitSetup& fs, FjErrorBuffer* error) {
    uint32_t category = FJ_STATUS_UNSUPPORTED_INPUT;
    std::string text;
    FjStatus st = fit_precheck(nominal, th, width, height, error);
    if (st.category != FJ_STATUS_SUCCESS) return st;
    FftPlanDev plan_w, plan_h;
    if (!make_fft_plan(width, plan_w) || !make_fft_plan(height, plan_h) || fft_length(plan_w) > dev.max_line ||
        fft_length(plan_h) > dev.max_line)
        return refuse(error, FJ_STATUS_UNSUPPORTED_INPUT, "UnsupportedInput spatial DIR frame side above the LDS line limit");
    if (!build_fit_record(nominal, th, fs.fr, text)) return refuse(error, FJ_STATUS_UNSUPPORTED_INPUT, text);
    ++dev.frame_serial;
    PixelParams& pp = fs.args.pp;
    // The parameters' own tables (film curves with the compensated axes,
    // print spectra with the filtered illuminant, and the adjoints' tables)
    // go through the upload ring: asynchronous copies on the stream, no
    // allocation and no cache churn per parameter vector.
    const std::vector<float> film = packed_film_curves(fs.fr.rec);
    int k = 0;
    const std::vector<float> spec = packed_print_spectra(fs.fr.rec.print, k);
    const size_t o_spec = aligned64(film.size()), o_alpha = o_spec + aligned64(spec.size()),
                 o_daxis = o_alpha + aligned64(fs.fr.alpha.size()), floats = o_daxis + aligned64(fs.fr.daxis.size());
    if (floats > DeviceState::kFitSlotFloats || fs.fr.alpha.size() != static_cast<size_t>(4 * (k > 0 ? k : 1)))
        return refuse(error, FJ_STATUS_UNSUPPORTED_INPUT, "UnsupportedInput fit tables above the upload slot");
    float *host = nullptr, *device = nullptr;
    hipError_t e = fs.ring.acquire(dev, stream, host, device);
    if (e != hipSuccess) return hip_fail(error, e, "fit upload ring");
    std::memcpy(host, film.data(), film.size() * sizeof(float));
    std::memcpy(host + o_spec, spec.data(), spec.size() * sizeof(float));
    std::memcpy(host + o_alpha, fs.fr.alpha.data(), fs.fr.alpha.size() * sizeof(float));
    std::memcpy(host + o_daxis, fs.fr.daxis.data(), fs.fr.daxis.size() * sizeof(float));
    e = hipMemcpyAsync(device, host, floats * sizeof(float), hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return hip_fail(error, e, "fit table upload");
    GivenTables given;
    given.film_curves = device;
    given.print_spectra = reinterpret_cast<const float4*>(device + o_spec);
    given.print_k = k;
    bool ok = false;
    e = assemble_pixel_params(dev, fs.fr.rec, pp, category, text, ok, &given);
    if (e != hipSuccess) return hip_fail(error, e, "fit table upload");
    if (!ok) return refuse(error, category, text);
    pp.io.width = width;
    pp.io.height = height;
    pp.io.comps = 4;
    pp.io.src_row_bytes = pp.io.dst_row_bytes = static_cast<long long>(width) * 16;
    fs.total = static_cast<long long>(width) * height;
    fs.args.alpha = reinterpret_cast<const float4*>(device + o_alpha);
    fs.args.daxis = device + o_daxis;
    std::memcpy(fs.args.mallett_A, fs.fr.A, sizeof(fs.fr.A));
    std::memcpy(fs.args.donor_k_unit, fs.fr.k_unit, sizeof(fs.fr.k_unit));
    fs.args.print_gamma = th.print_gamma;
    fs.args.film_gamma = th.film_gamma;
    fs.args.print_route_scale = nominal.scanner_correction.exposure_scale;
    fs.args.partial_blocks = fit_blocks(fs.total);
    // The spatial DIR operator at this frame size.
    bool tok = false;
    e = dir_transfer_cached(dev, fs.fr.rec.spatial_dir, width, height, fs.job.transfer, text, tok);
    if (e != hipSuccess) return hip_fail(error, e, "DIR transfer upload");
    if (!tok) return refuse(error, FJ_STATUS_UNSUPPORTED_INPUT, "MalformedPreparedHostData field=" + text);
    fs.job.width = width;
    fs.job.height = height;
    fs.job.planes = 3;
    return make_status(FJ_STATUS_SUCCESS);
}

// Holds the device workspace slot from acquire to release on `stream`.
struct SlotHold {
    DeviceState& dev;
    hipStream_t stream;
    bool held = false;
    hipError_t acquire(size_t bytes) {
        hipError_t e = dev.ensure_workspace(bytes);
        if (e != hipSuccess) return e;
        e = dev.acquire_workspace(stream);
        held = e == hipSuccess;
        return e;
    }
    hipError_t release() {
        if (!held) return hipSuccess;
        held = false;
        return dev.release_workspace(stream);
    }
    ~SlotHold() { (void)release(); }
};

}  // namespace

FjStatus fit_forward(DeviceState& dev, const FjPreparedHostData& nominal, const FjFitParams& th, const float* source,
                     int width, int height, const float* delta, float* destination, float* saved, hipStream_t stream,
                     FjErrorBuffer* error) {
    if (!source || !destination || !saved)
        return refuse(error, FJ_STATUS_UNSUPPORTED_INPUT, "UnsupportedInput fit null buffer");
    FitSetup fs;
    FjStatus st = fit_setup(dev, nominal, th, width, height, stream, fs, error);
    if (st.category != FJ_STATUS_SUCCESS) return st;
    const long long N = fs.total;
    PixelParams& pp = fs.args.pp;
    pp.io.src = reinterpret_cast<const char*>(source);
    pp.io.dst = reinterpret_cast<char*>(destination);
    for (int c = 0; c < 3; ++c) {
        pp.dir_planes[c] = saved + (3 + c) * N;  // raw, then filtered corrections
        pp.dir_planes[3 + c] = saved + c * N;    // log raw exposure
        pp.dens[c] = saved + (6 + c) * N;        // developed densities
    }
    pp.error_word = nullptr;
    SlotHold slot{dev, stream};
    hipError_t e = slot.acquire(static_cast<size_t>(6 * N) * sizeof(float));
    if (e != hipSuccess) return hip_fail(error, e, "workspace");
    float* ws = static_cast<float*>(dev.workspace());
    e = launch_pixel_stage(pp, kDirSource, dev.compute_units, stream);
    if (e != hipSuccess) return hip_fail(error, e, "fit DIR source");
    for (int c = 0; c < 3; ++c) {
        fs.job.in[c] = fs.job.out[c] = pp.dir_planes[c];
        fs.job.tmp_a[c] = ws + c * N;
        fs.job.tmp_b[c] = ws + (3 + c) * N;
    }
    e = run_dct_filter(dev, fs.job, stream);
    if (e != hipSuccess) return hip_fail(error, e, "fit DIR filter");
    e = launch_pixel_stage(pp, kDirFinal, dev.compute_units, stream);
    if (e != hipSuccess) return hip_fail(error, e, "fit DIR final");
    fs.args.src = source;
    fs.args.dst = destination;
    fs.args.delta = delta;
    fs.args.saved = saved;
    e = launch_fit_print(fs.args, stream);
    if (e != hipSuccess) return hip_fail(error, e, "fit print");
    e = slot.release();
    if (e != hipSuccess) return hip_fail(error, e, "workspace order");
    write_error(error, "");
    return make_status(FJ_STATUS_SUCCESS);
}

FjStatus fit_backward(DeviceState& dev, const FjPreparedHostData& nominal, const FjFitParams& th, const float* source,
                      int width, int height, const float* saved, const float* grad_destination, float* grad_source,
                      float* grad_params, hipStream_t stream, FjErrorBuffer* error) {
    if (!source || !saved || !grad_destination || !grad_source || !grad_params)
        return refuse(error, FJ_STATUS_UNSUPPORTED_INPUT, "UnsupportedInput fit null buffer");
    FitSetup fs;
    FjStatus st = fit_setup(dev, nominal, th, width, height, stream, fs, error);
    if (st.category != FJ_STATUS_SUCCESS) return st;
    const long long N = fs.total;
    const int blocks = fs.args.partial_blocks;
    // Workspace: gradient planes (density 3, corrections 3, log raw 3), DCT
    // scratch (6) and the partial sums.
    const size_t floats = static_cast<size_t>(15 * N) + static_cast<size_t>(3) * blocks * kFitSlots;
    SlotHold slot{dev, stream};
    hipError_t e = slot.acquire(floats * sizeof(float));
    if (e != hipSuccess) return hip_fail(error, e, "workspace");
    float* ws = static_cast<float*>(dev.workspace());
    FitArgs& a = fs.args;
    a.src = source;
    a.saved = const_cast<float*>(saved);
    a.grad_dst = grad_destination;
    a.grad_src = grad_source;
    a.g_dens = ws;
    a.g_filtered = ws + 3 * N;
    a.g_lograw = ws + 6 * N;
    a.partials = ws + 15 * N;
    e = launch_fit_adjoint(a, 0, stream);
    if (e == hipSuccess) e = launch_fit_adjoint(a, 1, stream);
    if (e != hipSuccess) return hip_fail(error, e, "fit adjoint");
    for (int c = 0; c < 3; ++c) {
        fs.job.in[c] = fs.job.out[c] = a.g_filtered + c * N;
        fs.job.tmp_a[c] = ws + (9 + c) * N;
        fs.job.tmp_b[c] = ws + (12 + c) * N;
    }
    // The DIR operator is symmetric: its adjoint is the same filter.
    e = run_dct_filter(dev, fs.job, stream);
    if (e != hipSuccess) return hip_fail(error, e, "fit DIR filter adjoint");
    e = launch_fit_adjoint(a, 2, stream);
    if (e == hipSuccess) e = launch_fit_reduce(a.partials, 3 * blocks, grad_params, stream);
    if (e != hipSuccess) return hip_fail(error, e, "fit source adjoint");
    e = slot.release();
    if (e != hipSuccess) return hip_fail(error, e, "workspace order");
    write_error(error, "");
    return make_status(FJ_STATUS_SUCCESS);
}

}  // namespace fjx
