// Batch schedules (include/optics_gpu.h, "Batch schedule tuner"): the picks'
// lookup and application to a call's prepared problems, the tuner
// (coordinate descent over the caller's own batch, timed as the batch call
// runs it) and the picks' text file.  No reference counterpart: the
// reference plans every transform with rustfft (src/propagation/fft.rs:40-64).
#include <chrono>
#include <cstdlib>
#include <fstream>
#include <set>
#include <sstream>

#include "runtime.hpp"

using namespace ocr;

namespace ocr {

uint64_t size_class(size_t n) {
    uint64_t c = 1;
    while (c < n) c <<= 1;
    return c;
}

static int log2_ceil(uint64_t c) {
    int l = 0;
    while ((1ull << l) < c) ++l;
    return l;
}

const oc_schedule* find_sched(const oc_context* ctx, bool pad, uint64_t len, bool dbl, uint64_t cls) {
    if (ctx->sched.empty()) return nullptr;
    const int kp = pad ? 0 : 1, kd = dbl ? 1 : 0;
    const oc_schedule* best = nullptr;
    int bd = 1 << 30;
    uint64_t bc = 0;
    for (auto it = ctx->sched.lower_bound(std::make_tuple(kp, len, kd, (uint64_t)0)); it != ctx->sched.end(); ++it) {
        if (std::get<0>(it->first) != kp || std::get<1>(it->first) != len || std::get<2>(it->first) != kd) break;
        const uint64_t c = std::get<3>(it->first);
        const int dist = std::abs(log2_ceil(c) - log2_ceil(cls));
        if (dist < bd || (dist == bd && c > bc)) {  // nearest class; the larger on a tie
            best = &it->second;
            bd = dist;
            bc = c;
        }
    }
    return best;
}

// Problems a schedule can steer: launched, square reference and run sizes.
static bool steerable(const Prob& p) {
    return p.status == OC_OK && p.launch && p.ref_nxp > 0 && p.ref_nxp == p.ref_nyp && p.nxp == p.nyp;
}

void apply_schedules(oc_context* ctx, bool dbl, std::vector<Prob>& probs) {
    if (ctx->sched.empty() || ctx->force_generic) return;
    std::map<size_t, size_t> nref, nrun;
    for (const Prob& p : probs)
        if (steerable(p)) ++nref[p.ref_nxp];
    for (Prob& p : probs) {
        if (!steerable(p)) continue;
        const oc_schedule* s = find_sched(ctx, true, p.ref_nxp, dbl, size_class(nref[p.ref_nxp]));
        // only the lengths the problem allows: the reference's, or the
        // power of two prepare() chose for a convolution
        if (s && (s->run_length == p.ref_nxp || s->run_length == p.nxp)) p.nxp = p.nyp = s->run_length;
    }
    for (const Prob& p : probs)
        if (steerable(p)) ++nrun[p.nxp];
    for (Prob& p : probs) {
        if (!steerable(p)) continue;
        const oc_schedule* s = find_sched(ctx, false, p.nxp, dbl, size_class(nrun[p.nxp]));
        if (!s) continue;
        if (s->kind == OC_SCHED_FAST) {
            if (s->variant >= 0 && s->variant < ocl::fast_variants((int)p.nxp)) {
                p.path = 1;
                p.variant = s->variant;
            }
        } else if (s->kind == OC_SCHED_LDS && (s->nradix == 0 || valid_radices((int)p.nxp, s->radix, s->nradix))) {
            p.path = 0;
            p.nrad = s->nradix;
            for (int i = 0; i < s->nradix; ++i) p.rad[i] = s->radix[i];
            if (!fits_on_chip(p, dbl)) {  // the LDS kernels must hold the length on chip
                p.path = -1;
                p.nrad = 0;
            }
        }
    }
}

}  // namespace ocr

namespace {

using Key = std::tuple<int, uint64_t, int, uint64_t>;

Key key_of(const oc_schedule& s) {
    return std::make_tuple(s.kind == OC_SCHED_PAD ? 0 : 1, s.length, s.double_precision ? 1 : 0,
                           size_class(s.problems));
}

bool sched_ok(const oc_schedule& s) {
    if (s.length == 0 || s.length > (1u << 20) || s.problems == 0) return false;
    switch (s.kind) {
        case OC_SCHED_PAD: return s.run_length >= s.length && s.run_length <= (1u << 20);
        case OC_SCHED_FAST: return s.variant >= 0 && s.variant < OC_FFT_VARIANTS;
        case OC_SCHED_LDS: return s.nradix == 0 || valid_radices((int)s.length, s.radix, s.nradix);
        default: return false;
    }
}

void install(oc_context* ctx, oc_schedule s) {
    s.problems = size_class(s.problems);
    s.double_precision = s.double_precision ? 1 : 0;
    ctx->sched[key_of(s)] = s;
}

// Radix plans of the in-place LDS stages for n = 2^a 3^b 5^c 7^d: the
// power-of-two part as radix 16 (the default), 8 or 4 stages plus one for
// the rest, 16s with a last 16 split into 4 x 4, or the default reversed;
// the odd factors after or before it.  Other lengths: none (their plan is
// the ping-pong one).
std::vector<std::vector<int>> radix_splits(int n) {
    int m = n, p2 = 0;
    while (m % 2 == 0) {
        m /= 2;
        ++p2;
    }
    std::vector<int> odd;
    for (int p : {3, 5, 7})
        while (m % p == 0) {
            odd.push_back(p);
            m /= p;
        }
    if (m != 1) return {};
    auto fill = [p2](int lg) {
        std::vector<int> r;
        int q = p2;
        while (q >= lg) {
            r.push_back(1 << lg);
            q -= lg;
        }
        if (q) r.push_back(1 << q);
        return r;
    };
    std::vector<std::vector<int>> pw{fill(4), fill(3), fill(2)};
    std::vector<int> r = fill(4);
    if (!r.empty() && r.back() == 16) {
        r.pop_back();
        r.push_back(4);
        r.push_back(4);
    }
    pw.push_back(r);
    r = fill(4);
    std::reverse(r.begin(), r.end());
    pw.push_back(r);
    std::set<std::vector<int>> seen;
    std::vector<std::vector<int>> out;
    for (const auto& a : pw)
        for (int first = 0; first < (odd.empty() ? 1 : 2); ++first) {
            std::vector<int> plan = first ? odd : a;
            plan.insert(plan.end(), first ? a.begin() : odd.begin(), first ? a.end() : odd.end());
            if (plan.empty() || (int)plan.size() > kMaxStages || !valid_radices(n, plan.data(), (int)plan.size()))
                continue;
            if (seen.insert(plan).second) out.push_back(plan);
        }
    return out;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int autotune(oc_context* ctx, bool dbl, oc_prop_desc* d, size_t count, int sweeps, oc_schedule* picks,
             size_t max_picks, size_t* npicks, double* search_ms) {
    if (!ctx || !d || count == 0) return OC_ERR_INVALID_ARGUMENT;
    if (ctx->arena) return fail(ctx, OC_ERR_INVALID_ARGUMENT, "the tuner allocates");
    (void)hipSetDevice(ctx->device);
    const double t0 = now_ms();
    std::vector<Prob> v;
    int st = probs_from_descs(d, count, v);
    if (st) return st;
    for (Prob& p : v) prepare(p);
    // reference lengths, the run lengths prepare() chose for them, counts
    std::map<size_t, std::set<size_t>> runs_of;
    std::map<size_t, size_t> nref;
    for (const Prob& p : v)
        if (steerable(p)) {
            runs_of[p.ref_nxp].insert(p.nxp);
            ++nref[p.ref_nxp];
        }
    struct Var {
        bool pad;
        uint64_t len;
        std::vector<oc_schedule> cands;  // cands[0]: the default
        size_t pick = 0;
    };
    std::vector<Var> vars;
    std::set<size_t> lengths;
    for (const auto& kv : runs_of) {
        lengths.insert(kv.first);
        for (size_t r : kv.second) lengths.insert(r);
        std::set<size_t> alt = kv.second;
        alt.erase(kv.first);
        if (alt.empty()) continue;
        Var x{true, kv.first, {}, 0};
        oc_schedule s{};
        s.length = kv.first;
        s.double_precision = dbl;
        s.kind = OC_SCHED_PAD;
        s.problems = size_class(nref[kv.first]);
        s.run_length = *alt.begin();  // prepare()'s choice first
        x.cands.push_back(s);
        s.run_length = kv.first;
        x.cands.push_back(s);
        vars.push_back(x);
    }
    for (size_t len : lengths) {
        Var x{false, len, {}, 0};
        oc_schedule s{};
        s.length = len;
        s.double_precision = dbl;
        s.problems = 1;  // set per evaluation
        const int nv = ocl::fast_variants((int)len);
        if (nv > 0) {
            s.kind = OC_SCHED_FAST;
            const int dv = ctx->variant_for((int)len, dbl);
            s.variant = dv;
            x.cands.push_back(s);
            for (int k = 0; k < nv; ++k)
                if (k != dv) {
                    s.variant = k;
                    x.cands.push_back(s);
                }
        }
        s.kind = OC_SCHED_LDS;
        s.variant = 0;
        s.nradix = 0;
        if (nv == 0) x.cands.push_back(s);  // the default plan first
        // each plan must hold every problem that can run at this length on chip
        for (const auto& plan : radix_splits((int)len)) {
            bool fits = true;
            for (const Prob& p : v) {
                if (!steerable(p) || (p.ref_nxp != len && !runs_of[p.ref_nxp].count(len))) continue;
                Prob q = p;
                q.nxp = q.nyp = len;
                fits = fits && fits_on_chip(q, dbl);
            }
            if (!fits) continue;
            s.nradix = (int)plan.size();
            for (size_t i = 0; i < plan.size(); ++i) s.radix[i] = plan[i];
            x.cands.push_back(s);
        }
        if (x.cands.size() > 1) vars.push_back(x);
    }
    // problems running at each length under the current pad picks
    auto run_count = [&](uint64_t len) {
        size_t c = 0;
        for (const auto& kv : nref) {
            uint64_t run = kv.first;
            for (const Var& x : vars)
                if (x.pad && x.len == kv.first) run = x.cands[x.pick].run_length;
            if (run != kv.first && !runs_of[kv.first].count(run)) run = kv.first;
            if (run == len) c += kv.second;
        }
        return c;
    };
    // the search owns every entry of its keys (all size classes): one pick
    // per key at a time
    auto is_var = [dbl](const Key& key, const Var& x) {
        return std::get<0>(key) == (x.pad ? 0 : 1) && std::get<1>(key) == x.len && std::get<2>(key) == (dbl ? 1 : 0);
    };
    auto install_var = [&](const Var& x, size_t c) {
        for (auto it = ctx->sched.begin(); it != ctx->sched.end();) it = is_var(it->first, x) ? ctx->sched.erase(it) : std::next(it);
        oc_schedule s = x.cands[c];
        if (!x.pad) s.problems = std::max<size_t>(1, run_count(x.len));
        install(ctx, s);
    };
    for (const Var& x : vars) install_var(x, 0);
    const bool timing = ctx->timing;
    ctx->timing = false;
    // the batch call's wall time, best of two after one untimed run; a
    // schedule whose statuses differ from the default's is refused
    std::vector<int> ref_status;
    int ref_ret = 0;
    auto time_batch = [&](double* ms, bool is_ref) -> int {
        *ms = 1e300;
        for (int rep = 0; rep < 3; ++rep) {
            std::vector<Prob> w;
            int s2 = probs_from_descs(d, count, w);
            if (s2) return s2;
            (void)hipStreamSynchronize(ctx->stream);
            const double a = now_ms();
            const int ret = run_batch(ctx, dbl, w);
            const double b = now_ms();
            if (ret == OC_ERR_DEVICE || ret == OC_ERR_OUT_OF_MEMORY) return ret;
            if (is_ref && rep == 0) {
                ref_ret = ret;
                ref_status.clear();
                for (const Prob& p : w) ref_status.push_back(p.status);
            } else {
                bool same = ret == ref_ret;
                for (size_t i = 0; i < w.size() && same; ++i) same = w[i].status == ref_status[i];
                if (!same) {
                    *ms = 1e300;
                    return OC_OK;
                }
            }
            for (size_t i = 0; i < count; ++i) d[i].status = w[i].status;
            if (rep > 0) *ms = std::min(*ms, b - a);
        }
        return OC_OK;
    };
    double cur = 0.0;
    st = time_batch(&cur, true);
    if (!st) {
        const int rounds = sweeps > 0 ? sweeps : 1;
        for (int r = 0; r < rounds && !st; ++r)
            for (Var& x : vars) {
                size_t best = x.pick;
                double best_ms = cur;
                for (size_t c = 0; c < x.cands.size() && !st; ++c) {
                    if (c == x.pick) continue;
                    install_var(x, c);
                    double ms;
                    st = time_batch(&ms, false);
                    if (!st && ms < best_ms) {
                        best_ms = ms;
                        best = c;
                    }
                }
                x.pick = best;
                cur = best_ms;
                install_var(x, best);
                x.cands[best].ms = best_ms;
            }
    }
    ctx->timing = timing;
    if (st) return st;
    // final size classes under the final pad picks, and the report
    size_t k = 0;
    for (const Var& x : vars) {
        install_var(x, x.pick);
        for (auto& kv : ctx->sched)
            if (is_var(kv.first, x)) {
                kv.second.ms = x.cands[x.pick].ms;
                if (picks && k < max_picks) picks[k] = kv.second;
                ++k;
                break;
            }
    }
    if (npicks) *npicks = std::min(k, picks ? max_picks : k);
    if (search_ms) *search_ms = now_ms() - t0;
    return OC_OK;
}

std::string device_tag(const oc_context* ctx) {
    hipDeviceProp_t prop;
    std::string name = "unknown";
    if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess) {
        name = prop.gcnArchName;
        const size_t colon = name.find(':');
        if (colon != std::string::npos) name = name.substr(0, colon);
    }
    return name + " " + std::to_string(ctx->num_cu);
}

}  // namespace

extern "C" {

int oc_autotune_batch_dev_c64(oc_context* ctx, oc_prop_desc* descs, size_t count, int sweeps, oc_schedule* picks,
                              size_t max_picks, size_t* npicks, double* search_ms) {
    OC_CONTEXT_LOCK(ctx);
    return autotune(ctx, false, descs, count, sweeps, picks, max_picks, npicks, search_ms);
}

int oc_autotune_batch_dev_c128(oc_context* ctx, oc_prop_desc* descs, size_t count, int sweeps, oc_schedule* picks,
                               size_t max_picks, size_t* npicks, double* search_ms) {
    OC_CONTEXT_LOCK(ctx);
    return autotune(ctx, true, descs, count, sweeps, picks, max_picks, npicks, search_ms);
}

int oc_schedule_set(oc_context* ctx, const oc_schedule* s) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || !s) return OC_ERR_INVALID_ARGUMENT;
    if (!sched_ok(*s)) return fail(ctx, OC_ERR_INVALID_ARGUMENT, "malformed schedule");
    oc_schedule c = *s;
    c.ms = 0.0;
    install(ctx, c);
    return OC_OK;
}

int oc_schedule_get(oc_context* ctx, size_t i, oc_schedule* out) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || !out || i >= ctx->sched.size()) return OC_ERR_INVALID_ARGUMENT;
    auto it = ctx->sched.begin();
    std::advance(it, (long)i);
    *out = it->second;
    return OC_OK;
}

size_t oc_schedule_count(oc_context* ctx) {
    OC_CONTEXT_LOCK(ctx);
    return ctx ? ctx->sched.size() : 0;
}

int oc_schedule_clear(oc_context* ctx) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx) return OC_ERR_INVALID_ARGUMENT;
    ctx->sched.clear();
    return OC_OK;
}

// Text format, one entry a line after the header and the device line:
//   pad  <length> <double> <class> <ms> <run_length>
//   fast <length> <double> <class> <ms> <variant>
//   lds  <length> <double> <class> <ms> <nradix> <radix>...
int oc_schedule_save(oc_context* ctx, const char* path) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || !path) return OC_ERR_INVALID_ARGUMENT;
    std::ofstream f(path);
    if (!f) return fail(ctx, OC_ERR_INVALID_ARGUMENT, std::string("cannot write ") + path);
    f << "optics_gpu schedules 1\n";
    f << "device " << device_tag(ctx) << "\n";
    f.precision(17);
    for (const auto& kv : ctx->sched) {
        const oc_schedule& s = kv.second;
        const char* kind = s.kind == OC_SCHED_PAD ? "pad" : s.kind == OC_SCHED_FAST ? "fast" : "lds";
        f << kind << " " << s.length << " " << s.double_precision << " " << s.problems << " " << s.ms;
        if (s.kind == OC_SCHED_PAD) f << " " << s.run_length;
        else if (s.kind == OC_SCHED_FAST) f << " " << s.variant;
        else {
            f << " " << s.nradix;
            for (int i = 0; i < s.nradix; ++i) f << " " << s.radix[i];
        }
        f << "\n";
    }
    f.close();
    if (!f) return fail(ctx, OC_ERR_INVALID_ARGUMENT, std::string("cannot write ") + path);
    return OC_OK;
}

int oc_schedule_load(oc_context* ctx, const char* path, size_t* loaded) {
    OC_CONTEXT_LOCK(ctx);
    if (!ctx || !path) return OC_ERR_INVALID_ARGUMENT;
    if (loaded) *loaded = 0;
    std::ifstream f(path);
    if (!f) return fail(ctx, OC_ERR_INVALID_ARGUMENT, std::string("cannot read ") + path);
    std::string line;
    if (!std::getline(f, line) || line != "optics_gpu schedules 1")
        return fail(ctx, OC_ERR_INVALID_ARGUMENT, "not a schedule file");
    if (!std::getline(f, line) || line.rfind("device ", 0) != 0)
        return fail(ctx, OC_ERR_INVALID_ARGUMENT, "schedule file without a device line");
    const bool mine = line.substr(7) == device_tag(ctx);
    std::vector<oc_schedule> entries;
    int ln = 2;
    while (std::getline(f, line)) {
        ++ln;
        if (line.empty()) continue;
        std::istringstream in(line);
        std::string kind;
        oc_schedule s{};
        in >> kind >> s.length >> s.double_precision >> s.problems >> s.ms;
        bool ok = (bool)in;
        if (kind == "pad") {
            s.kind = OC_SCHED_PAD;
            ok = ok && (bool)(in >> s.run_length);
        } else if (kind == "fast") {
            s.kind = OC_SCHED_FAST;
            ok = ok && (bool)(in >> s.variant);
        } else if (kind == "lds") {
            s.kind = OC_SCHED_LDS;
            ok = ok && (bool)(in >> s.nradix) && s.nradix >= 0 && s.nradix <= kMaxStages;
            for (int i = 0; ok && i < s.nradix; ++i) ok = (bool)(in >> s.radix[i]);
        } else {
            ok = false;
        }
        std::string rest;
        ok = ok && !(in >> rest) && sched_ok(s);
        if (!ok) return fail(ctx, OC_ERR_INVALID_ARGUMENT, "malformed schedule file, line " + std::to_string(ln));
        entries.push_back(s);
    }
    if (!mine) return OC_OK;  // tuned on another device: nothing applies here
    for (const oc_schedule& s : entries) {
        install(ctx, s);
        ctx->sched[key_of(s)].ms = s.ms;
    }
    if (loaded) *loaded = entries.size();
    return OC_OK;
}

}  // extern "C"
