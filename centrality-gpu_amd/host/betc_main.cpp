// betc: drop-in command line of the reference (Codes/main.cpp:54-72,
// Codes/ReadMe:3-4): `betc <edge_list_file>`, same input format, same stdout
// structure lines (Codes/modified_apsp.cpp:271,
// Codes/to_eargraph_construct.cpp:65, Codes/bcc_bfs.cpp:310,
// Codes/kernels_free.cu:10-11,113, Codes/begin_gpupgm.cu:69,
// Codes/kernels_active.cu:31,217,221, Codes/postprocesses.cu:200,
// Codes/main.cpp:50,70) and the same side-effect files (PostProcessOp.txt,
// Output.txt, parent.txt).  Timing values are measured by this program.
//
// Extra, opt-in flags:
//   --bc-out FILE        BC per vertex, "%d %f" (Codes/main.cpp:45-49)
//   --bc-out-exact FILE  BC per vertex, "%d %.17g"
//   --batched            multi-source batched path (default: plain): every
//                        component is uploaded once and run in one call, so
//                        the small components share one persistent launch
//   --fp32               FP32-compensated dependency arithmetic
//   --no-chains          every vertex a BFS source
//   --sample K           approximate BC: K sources per component (uniform,
//                        seeded; all sources when the component is smaller),
//                        each component's dependency sum scaled by n_c / K
//   --seed S             seed of --sample (default 1)
//   --cache              keep the reduced structure in <edge_list_file>.bcx
//                        and reuse it while the text file's size and
//                        modification time are unchanged (skips the text
//                        parse and the reduction on repeated runs)
#include <algorithm>
#include <charconv>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "bcx.h"

static int die(const char *what, const char *msg, int code) {
    std::fprintf(stderr, "betc: %s: %s\n", what, msg ? msg : "");
    return code;
}

// Sampled sources of a component (sorted local ids), or empty for all.
static std::vector<int32_t> sample_sources(int32_t n, int64_t k, uint64_t seed, int32_t comp) {
    std::vector<int32_t> s;
    if (k <= 0 || k >= n) return s;
    std::vector<int32_t> all((size_t)n);
    std::iota(all.begin(), all.end(), 0);
    std::mt19937_64 rng(seed * 0x9E3779B97F4A7C15ull + (uint64_t)comp);
    for (int64_t i = 0; i < k; ++i) {
        std::uniform_int_distribution<int64_t> d(i, n - 1);
        std::swap(all[(size_t)i], all[(size_t)d(rng)]);
    }
    s.assign(all.begin(), all.begin() + k);
    std::sort(s.begin(), s.end());
    return s;
}

// "%d %f\n" (or "%d %.17g\n") per vertex.  std::to_chars with a precision
// produces exactly what printf does for the same conversion; chunks of
// vertices are formatted on several threads and written in order.
static bool write_bc(FILE *f, const std::vector<double> &bc, bool exact) {
    const int64_t n = (int64_t)bc.size();
    const int64_t chunk = 1 << 16;
    const int64_t nch = (n + chunk - 1) / chunk;
    unsigned T = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::string> out((size_t)std::min<int64_t>(nch, 4 * (int64_t)T));
    for (int64_t c0 = 0; c0 < nch; c0 += (int64_t)out.size()) {
        const int64_t c1 = std::min<int64_t>(nch, c0 + (int64_t)out.size());
        auto fmt = [&](int64_t c) {
            std::string &s = out[(size_t)(c - c0)];
            s.clear();
            char buf[400];  // any double in %f fits
            for (int64_t v = c * chunk; v < std::min(n, (c + 1) * chunk); ++v) {
                char *p = std::to_chars(buf, buf + sizeof(buf), (int32_t)v).ptr;
                *p++ = ' ';
                p = exact ? std::to_chars(p, buf + sizeof(buf) - 1, bc[(size_t)v], std::chars_format::general, 17).ptr
                          : std::to_chars(p, buf + sizeof(buf) - 1, bc[(size_t)v], std::chars_format::fixed, 6).ptr;
                *p++ = '\n';
                s.append(buf, (size_t)(p - buf));
            }
        };
        std::vector<std::thread> th;
        for (unsigned t = 0; t < T; ++t)
            th.emplace_back([&, t] {
                for (int64_t c = c0 + t; c < c1; c += T) fmt(c);
            });
        for (auto &x : th) x.join();
        for (int64_t c = c0; c < c1; ++c) {
            const std::string &s = out[(size_t)(c - c0)];
            if (std::fwrite(s.data(), 1, s.size(), f) != s.size()) return false;
        }
    }
    return true;
}

int main(int argc, char **argv) {
    const char *in = nullptr, *bc_out = nullptr, *bc_exact = nullptr;
    bool batched = false, fp32 = false, chains = true, cache = false;
    int64_t sample = 0;
    uint64_t seed = 1;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--bc-out") && i + 1 < argc) bc_out = argv[++i];
        else if (!std::strcmp(argv[i], "--bc-out-exact") && i + 1 < argc) bc_exact = argv[++i];
        else if (!std::strcmp(argv[i], "--batched")) batched = true;
        else if (!std::strcmp(argv[i], "--fp32")) fp32 = true;
        else if (!std::strcmp(argv[i], "--no-chains")) chains = false;
        else if (!std::strcmp(argv[i], "--cache")) cache = true;
        else if (!std::strcmp(argv[i], "--sample") && i + 1 < argc) sample = std::atoll(argv[++i]);
        else if (!std::strcmp(argv[i], "--seed") && i + 1 < argc) seed = std::strtoull(argv[++i], nullptr, 10);
        else if (!in) in = argv[i];
    }
    if (!in) {
        std::printf("Usage : ./apsp INPUT_FILE_NAME1 INPUT_FILE_NAME2 .. [OUTPUT_FILE_NAME] [NUM_THREADS]\n");
        return 1;
    }
    bcx_host *h = nullptr;
    int rc = BCX_E_IO;
    const std::string cache_path = std::string(in) + ".bcx";
    if (cache && bcx_host_load(cache_path.c_str(), in, chains ? 1 : 0, &h) == BCX_OK) {
        std::fprintf(stderr, "betc: structure read from %s\n", cache_path.c_str());
    } else {
        rc = bcx_host_from_file(in, &h);
        if (rc == BCX_E_IO && !h) {
            std::printf("Input file cannot be opened\n");
            return 2;
        }
        if (rc) return die("read", "malformed input", 2);
        rc = bcx_host_decompose(h, chains ? 1 : 0, 1);
        if (rc) return die("decompose", bcx_host_last_error(h), 3);
        if (cache) {
            if (bcx_host_save(h, cache_path.c_str(), in) == BCX_OK)
                std::fprintf(stderr, "betc: structure written to %s\n", cache_path.c_str());
            else
                std::fprintf(stderr, "betc: cannot write %s (continuing without it)\n", cache_path.c_str());
        }
    }

    const int32_t n = bcx_host_num_vertices(h);
    const int32_t nb = bcx_host_num_bcc(h);
    std::printf(" n_bcc= %d\n", nb);
    bcx_ctx *ctx = nullptr;
    rc = bcx_create(&ctx, 0, 0);
    if (rc) return die("device", "no usable GPU", 4);
    int32_t n_cu = 0;
    bcx_device_info(ctx, &n_cu, nullptr, 0);

    bcx_opts o{};
    o.path = batched ? BCX_PATH_BATCHED : BCX_PATH_PLAIN;
    o.precision = fp32 ? BCX_PREC_FP32C : BCX_PREC_FP64;
    o.use_chains = chains ? 1 : 0;

    std::vector<double> bc((size_t)n, 0.0);
    {
        std::vector<int64_t> cut((size_t)n);
        bcx_host_cut_term(h, cut.data());
        for (int32_t v = 0; v < n; ++v) bc[v] = (double)cut[v];
    }
    double gpu_total = 0.0;
    int64_t active_total = 0, free_total = 0;
    const int32_t nc = bcx_host_num_components(h);
    const unsigned hw = std::thread::hardware_concurrency();
    struct Comp {
        bcx_component view;
        const int32_t *gid = nullptr;
        int32_t bcc = 0, nf = 0, ne = 0, nx = 0;
        int64_t offset = 0;
        std::vector<int32_t> smp;  // sampled sources (empty: all)
    };
    std::vector<Comp> cs((size_t)nc);
    int64_t total_n = 0;
    for (int32_t c = 0; c < nc; ++c) {
        Comp &x = cs[c];
        bcx_host_component(h, c, &x.view, &x.gid, &x.bcc);
        bcx_host_component_counts(h, c, &x.nf, &x.ne, &x.nx);
        x.offset = total_n;
        total_n += x.view.n;
        x.smp = sample_sources(x.view.n, sample, seed, c);
    }
    // the component's dependency sums, scaled for sampling, into bc
    auto scatter = [&](const Comp &x, const double *local) {
        const double f = x.smp.empty() ? 1.0 : (double)x.view.n / (double)x.smp.size();
        for (int32_t i = 0; i < x.view.n; ++i) bc[x.gid[i]] += f * local[i];
    };
    // reference stdout lines of one component (Codes/to_eargraph_construct.cpp:65,
    // Codes/bcc_bfs.cpp:310, Codes/kernels_free.cu:10-11,113,
    // Codes/begin_gpupgm.cu:69, Codes/postprocesses.cu:200,
    // Codes/kernels_active.cu:217,221) and its side-effect files
    auto print_comp = [&](int32_t c, double t_free, double t_post, double t_act) {
        const Comp &x = cs[c];
        const bcx_component &comp = x.view;
        std::printf("BCC number #: %d nodes %d edges %lld Total BCC's are %d\n", x.bcc, comp.n,
                    (long long)comp.arcs, nb);
        // the reference pins OpenMP to one thread after the first component
        std::printf("Num max threads %u\n", c == 0 ? (hw ? hw : 1) : 1);
        free_total += x.nf;
        if (x.nf > 0) {
            std::printf("Inside bc_gpu_free \n");
            std::printf(" Free vertices count is %d\n", x.nf);
            std::printf("total time for free vertices is ==================== %f\n", t_free);
            std::printf("Free vertices processing finished\n");
        }
        if (comp.n_chains > 0) {
            std::printf("max_threads_per_block %d number_of_SMs %d\n", comp.n < 1024 ? comp.n : 1024, n_cu);
            // the reference postprocesses every endpoint batch after the first
            // (Codes/kernels_active.cu:98-158), batches holding fewer than
            // 2000 endpoints (Codes/arrangelevels.cpp:8,84-109): with fewer
            // endpoints there is one batch and no such line.  (With more, the
            // reference's batch count follows its ear-graph BFS levels; one
            // line is printed here.)
            if (x.ne >= 2000) std::printf("postprocess time is %f\n", t_post);
            std::printf("total time for active vertices is %f\n", t_act);
            std::printf("total_act_vert_size is %d\n", x.ne);
            active_total += x.ne;
        }
    };
    if (batched) {
        // one upload and one run over every component; the run's kernel time
        // is reported against each component in proportion to its work
        std::vector<bcx_component> views;
        for (auto &x : cs) views.push_back(x.view);
        std::vector<int64_t> slots;
        for (auto &x : cs) {
            if (x.smp.empty()) for (int32_t v = 0; v < x.view.n; ++v) slots.push_back(x.offset + v);
            else for (int32_t v : x.smp) slots.push_back(x.offset + v);
        }
        bcx_dgraph *g = nullptr;
        if (nc > 0 && (rc = bcx_upload(ctx, views.data(), nc, &g))) return die("upload", bcx_last_error(ctx), 5);
        std::vector<double> local((size_t)std::max<int64_t>(total_n, 1), 0.0);
        bcx_stats st{};
        if (nc > 0) {
            bcx_opts ob = o;
            ob.src_subset = slots.data();
            ob.n_src = (int64_t)slots.size();
            rc = bcx_run(ctx, g, &ob, local.data(), &st);
            if (rc) return die("gpu", bcx_last_error(ctx), 6);
            bcx_dgraph_free(ctx, g);
        }
        double work = 0.0;
        std::vector<double> wc((size_t)nc);
        for (int32_t c = 0; c < nc; ++c) {
            const double ns = cs[c].smp.empty() ? cs[c].view.n : (double)cs[c].smp.size();
            wc[c] = ns * (double)cs[c].view.arcs;
            work += wc[c];
        }
        for (int32_t c = 0; c < nc; ++c) {
            const double t = work > 0 ? st.t_kernel_s * wc[c] / work : 0.0;
            const double share = (cs[c].nf + cs[c].ne) > 0 ? (double)cs[c].nf / (cs[c].nf + cs[c].ne) : 1.0;
            print_comp(c, t * share, 0.0, t * (1.0 - share));
            scatter(cs[c], local.data() + cs[c].offset);
        }
        gpu_total = st.t_kernel_s;
    } else {
        for (int32_t c = 0; c < nc; ++c) {
            const Comp &x = cs[c];
            const bcx_component &comp = x.view;
            std::vector<double> local((size_t)comp.n, 0.0);
            bcx_dgraph *g = nullptr;
            rc = bcx_upload(ctx, &comp, 1, &g);
            if (rc) return die("upload", bcx_last_error(ctx), 5);
            double t_free = 0.0, t_post = 0.0, t_act = 0.0;
            auto run = [&](std::vector<int64_t> &srcs, double &t_main, double *t_chain) -> int {
                if (srcs.empty()) return BCX_OK;
                bcx_opts os = o;
                os.src_subset = srcs.data();
                os.n_src = (int64_t)srcs.size();
                bcx_stats st{};
                int r = bcx_run(ctx, g, &os, local.data(), &st);
                if (r) return r;
                if (t_chain) {
                    *t_chain = st.t_chain_s;
                    t_main = st.t_kernel_s - st.t_chain_s;
                } else {
                    t_main = st.t_kernel_s;
                }
                gpu_total += st.t_kernel_s;
                return BCX_OK;
            };
            if (!x.smp.empty()) {
                // sampled: one run over the sample
                std::vector<int64_t> ss(x.smp.begin(), x.smp.end());
                if ((rc = run(ss, t_free, &t_post))) return die("gpu", bcx_last_error(ctx), 6);
            } else {
                // free sources, then endpoints + chain vertices (the active path)
                std::vector<int32_t> role((size_t)comp.n, 0);
                for (int32_t k = 0; k < comp.n_chains; ++k) {
                    role[comp.chain_ends[2 * k]] = 2;
                    role[comp.chain_ends[2 * k + 1]] = 2;
                    for (int32_t j = comp.chain_ptr[k]; j < comp.chain_ptr[k + 1]; ++j) role[comp.chain_nodes[j]] = 1;
                }
                std::vector<int64_t> fr, act;
                for (int32_t v = 0; v < comp.n; ++v) (role[v] == 0 ? fr : act).push_back(v);
                if ((rc = run(fr, t_free, nullptr)) || (rc = run(act, t_act, &t_post)))
                    return die("gpu", bcx_last_error(ctx), 6);
            }
            bcx_dgraph_free(ctx, g);
            print_comp(c, t_free, t_post, t_act);
            scatter(x, local.data());
        }
    }
    // side-effect files: the reference truncates and rewrites them for every
    // component (Codes/bcc_bfs.cpp:117,145-214), so what remains is the last
    // component's chain listing; written once here with that content
    if (nc > 0) {
        const bcx_component &comp = cs[nc - 1].view;
        if (FILE *pp = std::fopen("PostProcessOp.txt", "w")) {
            for (int32_t k = 0; k < comp.n_chains; ++k) {
                int32_t p0 = comp.chain_ptr[k], cnt = comp.chain_ptr[k + 1] - p0;
                std::fprintf(pp, "-1 %d %d %d %d\n", comp.chain_ends[2 * k], comp.chain_ends[2 * k + 1], cnt,
                             cnt + 1);
                for (int32_t j = 0; j < cnt; ++j)
                    std::fprintf(pp, "%d %d %d\n", j + 1, comp.chain_nodes[p0 + j], cnt - j);
            }
            std::fclose(pp);
        }
        if (FILE *f = std::fopen("Output.txt", "w")) std::fclose(f);
    }
    if (FILE *f = std::fopen("parent.txt", "w")) std::fclose(f);
    std::printf("active_vert_count %lld free_vert_count %lld \n", (long long)active_total, (long long)free_total);
    std::printf("GPU total time is %f \n", gpu_total);
    for (int k = 0; k < 2; ++k) {
        const char *path = k ? bc_exact : bc_out;
        if (!path) continue;
        FILE *f = std::fopen(path, "w");
        if (!f) return die("write", path, 7);
        if (!write_bc(f, bc, k == 1)) {
            std::fclose(f);
            return die("write", path, 7);
        }
        if (std::fclose(f) != 0) return die("write", path, 7);
    }
    bcx_destroy(ctx);
    bcx_host_free(h);
    return 0;
}
