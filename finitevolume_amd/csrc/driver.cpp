// Host side of the driver contract around the device solver: the FVMCKPT
// checkpoint format, the CFL ramp, and the per-step run loop with its
// residual CSV (src/main.cpp run_diffusion / run_euler / run_navier_stokes /
// run_ransSA / run_ransSST).  No kernels here: every device access goes
// through the solver's C entry points.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "solver.h"

namespace {

using fvk::set_error;

// Checkpoint header, include/Checkpoint.h:26-46 and src/Checkpoint.cpp
// (magic, format version 2, equation tag, cell count, step, build number;
// 40 bytes with no padding)
struct CkptHeader {
    char magic[8];
    uint32_t version;
    uint32_t equation;
    uint64_t cell_count;
    int64_t step_index;
    uint64_t build_number;
};
static_assert(sizeof(CkptHeader) == 40, "checkpoint header is 40 bytes");
constexpr char kMagic[8] = {'F', 'V', 'M', 'C', 'K', 'P', 'T', '\0'};
constexpr uint32_t kVersion = 2;

int ckpt_tag(int eq) {
    switch (eq) {
        case FVK_EQ_DIFFUSION: return FVK_CKPT_DIFFUSION;
        case FVK_EQ_ADVECTION_DIFFUSION: return FVK_CKPT_ADVECTION_DIFFUSION;
        case FVK_EQ_EULER: return FVK_CKPT_EULER;
        case FVK_EQ_NAVIER_STOKES: return FVK_CKPT_NAVIER_STOKES;
        case FVK_EQ_RANS_SA: return FVK_CKPT_RANS_SA;
        default: return FVK_CKPT_RANS_SST;
    }
}

bool scalar(int eq) { return eq == FVK_EQ_DIFFUSION || eq == FVK_EQ_ADVECTION_DIFFUSION; }

bool is_flow_field(uint32_t tag) {
    return tag == FVK_CKPT_EULER || tag == FVK_CKPT_NAVIER_STOKES || tag == FVK_CKPT_RANS_SA ||
           tag == FVK_CKPT_RANS_SST;
}

// std::ostream << double under std::setprecision(p), as the reference's CSVs
// are written
void put_row(std::ofstream& out, long long step, const double* r, int n) {
    out << step;
    for (int q = 0; q < n; ++q) out << "," << r[q];
    out << "\n";
    out.flush();
}

// is_nan_or_inf, src/main.cpp:65-69: the exponent bits all set
bool nonfinite(double x) {
    uint64_t b;
    std::memcpy(&b, &x, sizeof b);
    return ((b >> 52) & 0x7FF) == 0x7FF;
}

}  // namespace

extern "C" {

double fvk_next_cfl(double current_cfl, double cfl_max, const double* history, int32_t n_history, double cfl_min,
                    double growth_factor, double shrink_factor, int32_t window, double divergence_threshold) {
    // next_cfl, src/CflRamp.cpp:7-27: hold until the window is full, then
    // reset on a rise above the divergence threshold, grow on a falling
    // trend, shrink on a rise above the 10% dead band; clamp
    if (n_history < window || n_history <= 0 || !history) return current_cfl;
    const double trend = std::log10(history[n_history - 1]) - std::log10(history[0]);
    const double hold = std::log10(1.1);
    double next = current_cfl;
    if (trend > divergence_threshold) next = cfl_min;
    else if (trend < 0.0) next = current_cfl * growth_factor;
    else if (trend > hold) next = current_cfl * shrink_factor;
    // std::clamp(next, cfl_min, cfl_max)
    return next < cfl_min ? cfl_min : (cfl_max < next ? cfl_max : next);
}

int fvk_checkpoint_write(fvk_solver* s, const char* path, int64_t step_index, uint64_t build_number) {
    if (!s || !path) return set_error(FVK_ERR_INVALID, "NULL argument");
    const int64_t n = s->nc;
    const int eq = s->p.equation;
    const int nt = s->nv > 4 ? s->nv - 4 : 0;
    std::vector<double> U((scalar(eq) ? 1 : 4) * (size_t)n), T((size_t)nt * n);
    int rc = fvk_get_field(s, U.data(), nt ? T.data() : nullptr, n);
    if (rc != FVK_OK) return rc;
    std::ofstream out(path, std::ios::binary);
    if (!out.is_open()) return set_error(FVK_ERR_INVALID, "cannot open checkpoint file '%s' for writing", path);
    CkptHeader h;
    std::memcpy(h.magic, kMagic, sizeof kMagic);
    h.version = kVersion;
    h.equation = (uint32_t)ckpt_tag(eq);
    h.cell_count = (uint64_t)n;
    h.step_index = step_index;
    h.build_number = build_number;
    out.write(reinterpret_cast<const char*>(&h), sizeof h);
    // the EulerState payload (AoS), then nut, or k then omega, one array each
    out.write(reinterpret_cast<const char*>(U.data()), (std::streamsize)(sizeof(double) * U.size()));
    for (int q = 0; q < nt; ++q) {
        std::vector<double> col((size_t)n);
        for (int64_t c = 0; c < n; ++c) col[c] = T[(size_t)c * nt + q];
        out.write(reinterpret_cast<const char*>(col.data()), (std::streamsize)(sizeof(double) * n));
    }
    if (!out) return set_error(FVK_ERR_INVALID, "writing checkpoint file '%s' failed", path);
    return FVK_OK;
}

int fvk_checkpoint_read(fvk_solver* s, const char* path, int64_t* step_index, uint64_t* build_number,
                        int32_t* source_equation) {
    if (!s || !path) return set_error(FVK_ERR_INVALID, "NULL argument");
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) return set_error(FVK_ERR_INVALID, "cannot open checkpoint file '%s'", path);
    CkptHeader h;
    in.read(reinterpret_cast<char*>(&h), sizeof h);
    if (!in) return set_error(FVK_ERR_INVALID, "checkpoint file '%s' is truncated or corrupt (short header)", path);
    if (std::memcmp(h.magic, kMagic, sizeof kMagic) != 0)
        return set_error(FVK_ERR_INVALID, "checkpoint file '%s' is not a valid FiniteVolume checkpoint (bad magic)",
                         path);
    if (h.version != kVersion)
        return set_error(FVK_ERR_INVALID, "checkpoint file '%s' has unsupported version %u (expected %u)", path,
                         h.version, kVersion);
    const int eq = s->p.equation;
    const int64_t n = s->nc;
    // the resume rules of the reference driver: the scalar solvers need
    // their own tag (src/main.cpp:3117, 3242); the flow solvers take any flow field
    // (read_flow_field, :3461, 3662, 3940, 4260) and the turbulence fields
    // only from a checkpoint of their own model (:3944-3960, 4264-4281)
    if (scalar(eq) ? h.equation != (uint32_t)ckpt_tag(eq) : !is_flow_field(h.equation))
        return set_error(FVK_ERR_INVALID, "checkpoint file '%s' holds equation tag %u, which this solver cannot resume",
                         path, h.equation);
    if (h.cell_count != (uint64_t)n)
        return set_error(FVK_ERR_INVALID, "checkpoint file '%s' has %llu cells but the mesh has %lld", path,
                         (unsigned long long)h.cell_count, (long long)n);
    const int nw = scalar(eq) ? 1 : 4;
    std::vector<double> U((size_t)nw * n);
    in.read(reinterpret_cast<char*>(U.data()), (std::streamsize)(sizeof(double) * U.size()));
    if (!in) return set_error(FVK_ERR_INVALID, "checkpoint file '%s' is truncated or corrupt (short payload)", path);
    int nt = 0;
    if (eq == FVK_EQ_RANS_SA && h.equation == FVK_CKPT_RANS_SA) nt = 1;
    if (eq == FVK_EQ_RANS_SST && h.equation == FVK_CKPT_RANS_SST) nt = 2;
    std::vector<double> T((size_t)nt * n);
    for (int q = 0; q < nt; ++q) {
        std::vector<double> col((size_t)n);
        in.read(reinterpret_cast<char*>(col.data()), (std::streamsize)(sizeof(double) * n));
        if (!in) return set_error(FVK_ERR_INVALID, "checkpoint file '%s' is truncated or corrupt (short payload)", path);
        for (int64_t c = 0; c < n; ++c) T[(size_t)c * nt + q] = col[c];
    }
    int rc = fvk_set_field(s, U.data(), nt ? T.data() : nullptr, n);
    if (rc != FVK_OK) return rc;
    if (step_index) *step_index = h.step_index;
    if (build_number) *build_number = h.build_number;
    if (source_equation) *source_equation = (int32_t)h.equation;
    return FVK_OK;
}

void fvk_run_config_default(fvk_run_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof *c);
    c->nsteps = 0;
    c->cfl_ramp = 0;
    c->ramp_window = 15;
    c->cfl = 0.5;
    c->ramp_cfl_min = 0.05;
    c->ramp_growth = 1.5;
    c->ramp_shrink = 0.5;
    c->ramp_divergence_threshold = 2.0;
    for (double& t : c->tolerance) t = -1.0;
    c->residual_interval = 1;
    c->output_precision = 6;
}

int fvk_run(fvk_solver* s, const fvk_run_config* cfg, fvk_run_result* res) {
    fvk::TraceRange tr("fvk_run");
    if (!s || !cfg || !res) return set_error(FVK_ERR_INVALID, "NULL argument");
    if (cfg->residual_interval <= 0 || cfg->output_precision < 1 || cfg->output_precision > 17 ||
        (cfg->cfl_ramp && cfg->ramp_window <= 0))
        return set_error(FVK_ERR_INVALID, "bad run configuration");
    const int eq = s->p.equation;
    const int nv = s->nv;
    std::memset(res, 0, sizeof *res);
    res->resumed_equation = -1;
    // ramp start (src/main.cpp:3438-3442): from cfl_min, not the ceiling
    double cfl = cfg->cfl;
    const bool ramp = cfg->cfl_ramp && !scalar(eq);
    if (ramp) cfl = cfg->ramp_cfl_min;
    if (!scalar(eq)) {
        if (int rc = fvk_set_cfl(s, cfl); rc != FVK_OK) return rc;
    }
    // auto-resume (src/main.cpp:3449-3472)
    const bool ckpt = cfg->checkpoint_file && cfg->checkpoint_file[0];
    long long resume_start = 1;
    if (ckpt) {
        std::ifstream probe(cfg->checkpoint_file);
        if (probe.is_open()) {
            probe.close();
            int64_t step = 0;
            uint64_t build = 0;
            int32_t src = -1;
            if (int rc = fvk_checkpoint_read(s, cfg->checkpoint_file, &step, &build, &src); rc != FVK_OK) return rc;
            resume_start = step + 1;
            res->resumed_equation = src;
        }
    }
    res->resume_start = resume_start;
    // residual CSV (src/main.cpp:3474-3480; NS :3676-3681, SA :3969-3974, SST :4290-4295)
    std::ofstream csv;
    const bool tracking = cfg->residual_file && cfg->residual_file[0];
    if (tracking) {
        csv.open(cfg->residual_file);
        if (!csv.is_open()) return set_error(FVK_ERR_INVALID, "cannot open residual file '%s'", cfg->residual_file);
        csv << std::setprecision(cfg->output_precision);
        switch (eq) {
            case FVK_EQ_DIFFUSION:
            case FVK_EQ_ADVECTION_DIFFUSION: csv << "step,residual\n"; break;
            case FVK_EQ_RANS_SA: csv << "step,residual_rho,residual_rho_u,residual_rho_v,residual_E,residual_nut\n"; break;
            case FVK_EQ_RANS_SST:
                csv << "step,residual_rho,residual_rho_u,residual_rho_v,residual_E,residual_k,residual_omega\n";
                break;
            default: csv << "step,residual_rho,residual_rho_u,residual_rho_v,residual_E\n";
        }
    }
    bool checking = false;
    for (int q = 0; q < nv; ++q) checking = checking || cfg->tolerance[q] >= 0.0;
    std::vector<double> window;
    long long last = resume_start - 1;
    int status = 0;
    double r[6] = {0, 0, 0, 0, 0, 0};
    // the step loop, src/main.cpp:3491-3534 (one step, its residual row, the
    // divergence test, the ramp, the CSV row, the convergence test)
    for (long long k = resume_start; k <= cfg->nsteps; ++k) {
        if (int rc = fvk_step(s, 1, r, nullptr); rc != FVK_OK) return rc;
        last = k;
        bool bad = false;
        for (int q = 0; q < nv; ++q) bad = bad || nonfinite(r[q]);
        if (bad) {
            status = 2;
            break;
        }
        if (ramp) {
            window.push_back(r[0]);
            if ((int)window.size() > cfg->ramp_window) window.erase(window.begin());
            cfl = fvk_next_cfl(cfl, cfg->cfl, window.data(), (int32_t)window.size(), cfg->ramp_cfl_min,
                               cfg->ramp_growth, cfg->ramp_shrink, cfg->ramp_window, cfg->ramp_divergence_threshold);
            if (int rc = fvk_set_cfl(s, cfl); rc != FVK_OK) return rc;
        }
        if (tracking && k % cfg->residual_interval == 0) put_row(csv, k, r, nv);
        if (checking) {
            bool ok = true;
            for (int q = 0; q < nv; ++q) ok = ok && (cfg->tolerance[q] < 0.0 || r[q] < cfg->tolerance[q]);
            if (ok) {
                status = 1;
                break;
            }
        }
    }
    res->last_step = last;
    res->status = status;
    std::memcpy(res->residual, r, sizeof r);
    res->cfl = cfl;
    // the closing checkpoint (src/main.cpp:3551-3557): not after a divergence
    if (status != 2 && ckpt && last >= resume_start) {
        if (int rc = fvk_checkpoint_write(s, cfg->checkpoint_file, last, cfg->build_number); rc != FVK_OK) return rc;
    }
    return fvk_synchronize(s);
}

}  // extern "C"
