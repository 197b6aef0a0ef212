// abm_main.cpp -- the `abm` command-line driver on top of libparagraph.so.
//
// Drop-in for the reference's `abm` executable (src/main.cu:22-175): the same
// flags (src/main.cu:29-114; booleans accept true/1/yes, src/main.cu:43-45),
// the same inputs -- nodelist `node,year` remapped to dense ids in ascending
// original-id order (src/graph.cu:336-373), edgelist `citing,cited` with
// unknown ids dropped (src/graph.cu:376-447), delimiter detected from the
// first line as ',', tab, then space (includes/graph.cuh:41-53), `#` lines
// skipped -- and the same outputs: `#source,target` edgelist rows in source
// order with targets ascending (src/graph.cu:909-1003) and the 15-column
// auxiliary CSV (src/graph.cu:1057-1218).  New papers get original ids
// max_id+1, ... (src/kernel.cu:1638, 1711-1719), and each year adds
// ceil(N * growth_rate) papers (src/abm.cu:233-240).
//
// Models: `--model pa` (the default) and `er-gnp` are this library's PA/ER
// growth models (weights (k+1)^a * exp(-age/tau) * fitness, FP64 CDF, 2-hop
// neighbourhood draws, or ER G(n,p)); with them the flags of the reference's
// score-mixing model are parsed and reported as not used.  `--model
// reference` runs the reference's scoring model itself (PG_MODEL_ABM): the
// out-degree bag (src/abm.cu:70-88, values above 250 clamped), the recency
// table (src/abm.cu:91-110), --alpha (negative: per-paper uniform,
// src/abm.cu:241-254), --preferential/recency/fitness-weight (-1: per-paper
// uniform, src/abm.cu:256-280), --fully-random-citations and
// --same-year-proportion, with the reference's defaults (src/main.cu:61-85);
// its auxiliary CSV carries the reference's column meanings.  Planted nodes
// (--planted-nodes, src/abm.cu:44-68) are not modelled.
// The process exits non-zero on any failure (the reference always exits 0,
// src/main.cu:169).
#include <algorithm>
#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <exception>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/paragraph.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Args {
    std::map<std::string, std::string> kv;
    bool has(const std::string& k) const { return kv.count(k) != 0; }
    std::string str(const std::string& k, const std::string& def = "") const {
        auto it = kv.find(k);
        return it == kv.end() ? def : it->second;
    }
    double num(const std::string& k, double def) const {
        auto it = kv.find(k);
        return it == kv.end() ? def : std::stod(it->second);
    }
    bool flag(const std::string& k, bool def) const {
        auto it = kv.find(k);
        if (it == kv.end()) return def;
        const std::string& v = it->second;
        return v == "true" || v == "1" || v == "yes";
    }
};

const char* kKnown[] = {
    // reference flags (src/main.cu:29-114)
    "--edgelist", "--nodelist", "--out-degree-bag", "--recency-probabilities", "--use-warp-bfs",
    "--use-batching", "--use-multistage-kernel", "--max-batch-size", "--planted-nodes",
    "--preferential-weight", "--recency-weight", "--fitness-weight", "--alpha",
    "--fully-random-citations", "--growth-rate", "--num-cycles", "--same-year-proportion",
    "--output-file", "--auxiliary-information-file", "--log-file", "--num-processors",
    "--log-level",
    // ER flags of the reference's gpu-er build (results/slurm-8640738.out:553-556)
    "--model", "--er-probability", "--er-edges-per-node",
    // this library's model parameters
    "--pa-exponent", "--age-decay", "--neighbourhood-fraction", "--citations-per-paper",
    "--seed", "--bfs-cap", "--fitness-sigma", "--device",
    // exact year-granular checkpoint / resume (no reference counterpart)
    "--checkpoint-file", "--resume-from",
    // `node,year` list of the grown graph, the nodelist input of a chained run
    "--output-nodelist"};

const char* kRequired[] = {"--edgelist", "--nodelist", "--out-degree-bag",
                           "--recency-probabilities", "--growth-rate", "--num-cycles",
                           "--same-year-proportion", "--output-file",
                           "--auxiliary-information-file", "--log-file"};

void usage() {
    std::fprintf(stderr,
                 "usage: abm --edgelist F --nodelist F --out-degree-bag F --recency-probabilities F\n"
                 "           --growth-rate R --num-cycles Y --same-year-proportion P\n"
                 "           --output-file F --auxiliary-information-file F --log-file F\n"
                 "           [--model pa|er-gnp|reference] [--er-probability p] [--er-edges-per-node k]\n"
                 "           [--pa-exponent a] [--age-decay tau] [--neighbourhood-fraction f | --alpha f]\n"
                 "           [--citations-per-paper m] [--seed s] [--bfs-cap c] [--fitness-sigma s]\n"
                 "           [--checkpoint-file F (written after every year)] [--resume-from F]\n"
                 "           [--output-nodelist F (node,year of every paper, for chaining runs)]\n");
}

Args parse(int argc, char** argv) {
    Args a;
    for (int i = 1; i < argc; ++i) {
        std::string k = argv[i];
        if (k == "-h" || k == "--help") { usage(); std::exit(0); }
        bool known = false;
        for (const char* kk : kKnown) known = known || k == kk;
        if (!known) throw std::runtime_error("unknown argument " + k);
        if (i + 1 >= argc) throw std::runtime_error("missing value for " + k);
        a.kv[k] = argv[++i];
    }
    for (const char* r : kRequired)
        if (!a.has(r)) throw std::runtime_error(std::string(r) + ": required.");
    return a;
}

char detect_delimiter(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::string line;
    std::getline(f, line);
    if (line.find(',') != std::string::npos) return ',';
    if (line.find('\t') != std::string::npos) return '\t';
    if (line.find(' ') != std::string::npos) return ' ';
    throw std::invalid_argument("Could not detect filetype for " + path);
}

std::string slurp(const std::string& path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::string s;
    s.resize((size_t)f.tellg());
    f.seekg(0);
    f.read(&s[0], (std::streamsize)s.size());
    return s;
}

// `key,value` lines of a small table (first field an integer, second a
// number), `#` lines skipped, delimiter detected as for the other inputs.
std::vector<std::pair<long long, double>> read_table(const std::string& path) {
    const char delim = detect_delimiter(path);
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::vector<std::pair<long long, double>> out;
    std::string line;
    while (std::getline(f, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty() || line[0] == '#') continue;
        const size_t c = line.find(delim);
        if (c == std::string::npos) throw std::runtime_error("malformed line in " + path + ": " + line);
        try {
            out.emplace_back(std::stoll(line.substr(0, c)), std::stod(line.substr(c + 1)));
        } catch (const std::exception&) {
            throw std::runtime_error("malformed line in " + path + ": " + line);
        }
    }
    return out;
}

// Host threads for parsing and formatting.
int io_threads() {
    const unsigned h = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(h, 16u));
}

// Runs fn(t) for t in [0, T) on T threads; rethrows the first failure in t
// order.
template <typename F>
void parallel_for(int T, F&& fn) {
    std::vector<std::exception_ptr> err(T);
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t)
        th.emplace_back([&, t] {
            try { fn(t); } catch (...) { err[t] = std::current_exception(); }
        });
    for (auto& x : th) x.join();
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
}

// Splits [p, end) into lines and the first two delimiter-separated fields.
template <typename F>
void for_each_pair(const char* p, const char* end, char delim, F&& fn) {
    while (p < end) {
        const char* le = p;
        while (le < end && *le != '\n' && *le != '\r') ++le;
        if (le > p && *p != '#') {
            const char* f1 = p;
            while (f1 < le && *f1 != delim) ++f1;
            const char* s2 = f1 < le ? f1 + 1 : f1;
            const char* f2 = s2;
            while (f2 < le && *f2 != delim) ++f2;
            long long a = 0, b = 0;
            auto r1 = std::from_chars(p, f1, a);
            auto r2 = std::from_chars(s2, f2, b);
            if (r1.ec != std::errc() || r2.ec != std::errc())
                throw std::runtime_error("malformed line: " + std::string(p, le));
            fn(a, b);
        }
        p = le;
        while (p < end && (*p == '\n' || *p == '\r')) ++p;
    }
}

// The (first, second) field pairs of every line, in file order: T chunks cut
// at line starts are parsed on T threads.
std::vector<std::pair<long long, long long>> read_pairs(const std::string& path, int T) {
    const char delim = detect_delimiter(path);
    const std::string text = slurp(path);
    std::vector<size_t> cut(T + 1, text.size());
    cut[0] = 0;
    for (int t = 1; t < T; ++t) {
        size_t q = std::max(cut[t - 1], text.size() / T * t);
        while (q < text.size() && q > 0 && text[q - 1] != '\n') ++q;
        cut[t] = q;
    }
    std::vector<std::vector<std::pair<long long, long long>>> part(T);
    parallel_for(T, [&](int t) {
        for_each_pair(text.data() + cut[t], text.data() + cut[t + 1], delim,
                      [&](long long a, long long b) { part[t].emplace_back(a, b); });
    });
    size_t tot = 0;
    for (auto& v : part) tot += v.size();
    std::vector<std::pair<long long, long long>> all;
    all.reserve(tot);
    for (auto& v : part) all.insert(all.end(), v.begin(), v.end());
    return all;
}

struct Logger {
    std::ofstream f;
    int level;
    Clock::time_point t0 = Clock::now();
    void info(const std::string& m) { write("[INFO]", m); }
    void debug(const std::string& m) { if (level >= 1) write("[DEBUG]", m); }
    void write(const char* tag, const std::string& m) {
        if (level < 0 || !f) return;
        double s = std::chrono::duration<double>(Clock::now() - t0).count();
        f << tag << " " << m << " (elapsed " << s << " s)\n";
        f.flush();
    }
};

void check(int rc, pg_ctx* ctx, const char* what) {
    if (rc != PG_OK)
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " +
                                 pg_last_error(ctx));
}

void append_int(std::string& buf, long long v) {
    char tmp[24];
    auto r = std::to_chars(tmp, tmp + sizeof(tmp), v);
    buf.append(tmp, r.ptr);
}

}  // namespace

int main(int argc, char** argv) {
    try {
        const Args a = parse(argc, argv);
        Logger log;
        log.level = (int)a.num("--log-level", 1) - 1;
        log.f.open(a.str("--log-file"));
        std::string model = a.str("--model", "pa");
        if (model == "abm") model = "reference";
        if (model != "pa" && model != "er-gnp" && model != "reference")
            throw std::runtime_error("--model must be pa, er-gnp or reference");
        const bool ref_model = model == "reference";
        const double rate = a.num("--growth-rate", 0.0);
        const int years = (int)a.num("--num-cycles", 0);
        if (!(rate >= 0.0) || years < 0) throw std::runtime_error("bad --growth-rate/--num-cycles");
        for (const char* k : {"--preferential-weight", "--recency-weight", "--fitness-weight",
                              "--same-year-proportion", "--fully-random-citations", "--planted-nodes",
                              "--use-warp-bfs", "--use-batching", "--max-batch-size",
                              "--use-multistage-kernel", "--num-processors"}) {
            const bool exec_knob = std::strcmp(k, "--use-warp-bfs") == 0 ||
                                   std::strcmp(k, "--use-batching") == 0 ||
                                   std::strcmp(k, "--max-batch-size") == 0 ||
                                   std::strcmp(k, "--use-multistage-kernel") == 0 ||
                                   std::strcmp(k, "--num-processors") == 0;
            if (!a.has(k)) continue;
            if (!ref_model)
                std::fprintf(stderr, "abm: %s is a parameter of the reference's score-mixing "
                                     "model and is not used by this growth model\n", k);
            else if (exec_knob)
                std::fprintf(stderr, "abm: %s selects the reference's kernel schedule and has no "
                                     "effect here\n", k);
        }
        if (ref_model && !a.str("--planted-nodes").empty())
            std::fprintf(stderr, "abm: --planted-nodes is not modelled; planted papers get drawn "
                                 "fitness like the others\n");

        const int T = io_threads();
        const auto t_read = Clock::now();
        // ---- nodelist: dense ids in ascending original-id order (a repeated
        // id keeps its last year) ----
        std::vector<long long> orig;
        std::vector<int32_t> birth;
        {
            auto nodes = read_pairs(a.str("--nodelist"), T);
            std::stable_sort(nodes.begin(), nodes.end(),
                             [](const auto& x, const auto& y) { return x.first < y.first; });
            for (size_t i = 0; i < nodes.size(); ++i) {
                if (i + 1 < nodes.size() && nodes[i + 1].first == nodes[i].first) continue;
                orig.push_back(nodes[i].first);
                birth.push_back((int32_t)nodes[i].second);
            }
        }
        if (orig.empty()) throw std::runtime_error("empty nodelist");
        const int64_t n0 = (int64_t)orig.size();
        int max_year = birth[0];
        for (int32_t y : birth) max_year = std::max(max_year, y);
        // dense index of an original id, -1 if unknown (direct when the ids
        // form one range)
        const bool dense = orig[n0 - 1] - orig[0] + 1 == n0;
        auto index_of = [&](long long id) -> int32_t {
            if (dense) return (id >= orig[0] && id <= orig[n0 - 1]) ? (int32_t)(id - orig[0]) : -1;
            auto it = std::lower_bound(orig.begin(), orig.end(), id);
            return (it != orig.end() && *it == id) ? (int32_t)(it - orig.begin()) : -1;
        };
        // ---- edgelist: edges between known papers, in file order ----
        std::vector<int32_t> src, dst;
        {
            const auto edges = read_pairs(a.str("--edgelist"), T);
            std::vector<std::vector<int32_t>> ps(T), pd(T);
            parallel_for(T, [&](int t) {
                const size_t b = edges.size() / T * t;
                const size_t e = t + 1 == T ? edges.size() : edges.size() / T * (t + 1);
                for (size_t k = b; k < e; ++k) {
                    const int32_t iu = index_of(edges[k].first), iv = index_of(edges[k].second);
                    if (iu >= 0 && iv >= 0 && iu != iv) {
                        ps[t].push_back(iu);
                        pd[t].push_back(iv);
                    }
                }
            });
            for (int t = 0; t < T; ++t) {
                src.insert(src.end(), ps[t].begin(), ps[t].end());
                dst.insert(dst.end(), pd[t].begin(), pd[t].end());
            }
        }
        const double read_s = std::chrono::duration<double>(Clock::now() - t_read).count();
        // ---- citations per new paper (out-degree bag `idx,outdeg`) ----
        int m = (int)a.num("--citations-per-paper", -1);
        // ---- the reference's scoring model: bag and recency table ----
        std::vector<int32_t> bag;
        std::vector<double> rec;
        int32_t max_out = 0;
        if (ref_model) {
            const std::string path = a.str("--out-degree-bag");
            const std::string text = slurp(path);
            int64_t clamped = 0;
            for_each_pair(text.data(), text.data() + text.size(), detect_delimiter(path),
                          [&](long long, long long k) {
                              if (k < 0) throw std::runtime_error("negative out-degree in the bag");
                              if (k > 250) ++clamped;
                              bag.push_back((int32_t)std::min<long long>(k, 250));
                          });
            if (bag.empty()) throw std::runtime_error("empty out-degree bag");
            if (clamped)
                std::fprintf(stderr, "abm: %lld out-degree bag values above 250 (the reference's "
                                     "citation buffer) clamped to 250\n", (long long)clamped);
            for (int32_t k : bag) max_out = std::max(max_out, k);
            for (const auto& [d, prob] : read_table(a.str("--recency-probabilities"))) {
                if (d < 0 || d >= 1024)
                    throw std::runtime_error("recency table: year differences must be in [0, 1023]");
                if (!(prob >= 0.0)) throw std::runtime_error("recency table: negative probability");
                if ((size_t)d >= rec.size()) rec.resize((size_t)d + 1, 0.0);
                rec[(size_t)d] = prob;  // a repeated difference keeps its last value
            }
            if (a.num("--alpha", -1.0) > 1.0) throw std::runtime_error("--alpha must be at most 1");
            m = 0;
        }
        if (m < 0 && model == "pa") {
            std::vector<long long> bag;
            const std::string path = a.str("--out-degree-bag");
            const std::string text = slurp(path);
            for_each_pair(text.data(), text.data() + text.size(), detect_delimiter(path),
                          [&](long long, long long k) { bag.push_back(k); });
            if (bag.empty()) throw std::runtime_error("empty out-degree bag");
            for (long long k : bag)
                if (k != bag[0])
                    throw std::runtime_error("this growth model draws a fixed number of citations "
                                             "per paper; the out-degree bag holds several values, "
                                             "pass --citations-per-paper");
            m = (int)bag[0];
        }
        if (model == "er-gnp") m = 0;
        if (m > 64) throw std::runtime_error("at most 64 citations per paper");
        if (!ref_model && a.has("--alpha") && !a.has("--neighbourhood-fraction") &&
            !(a.num("--alpha", 0.0) >= 0.0))
            throw std::runtime_error("a negative --alpha (the reference's per-paper uniform "
                                     "neighbourhood share) is not supported by this growth "
                                     "model; pass a share in [0, 1]");
        const double f_nb = a.has("--neighbourhood-fraction") ? a.num("--neighbourhood-fraction", 0.3)
                            : a.num("--alpha", 0.3);
        int n_nb = (int)std::floor(f_nb * m + 0.5);
        if (m > 0 && n_nb > m - 1) n_nb = m - 1;
        if (n_nb < 0) n_nb = 0;
        const double alpha = a.num("--pa-exponent", 1.0);
        const double tau = a.num("--age-decay", 5.0);
        const uint64_t seed = (uint64_t)a.num("--seed", 1);
        double p_er = a.num("--er-probability", -1.0);
        // ---- schedule and capacities ----
        std::vector<int64_t> sched;
        int64_t n = n0, max_new = 0;
        for (int y = 0; y < years; ++y) {
            int64_t dn = pg_growth_count(n, rate);
            sched.push_back(dn);
            max_new = std::max(max_new, dn);
            n += dn;
        }
        const int64_t n_final = n;
        if (model == "er-gnp" && p_er < 0.0) {
            const double k = a.num("--er-edges-per-node", -1.0);
            if (k < 0.0) throw std::runtime_error("--model er-gnp needs --er-probability or --er-edges-per-node");
            p_er = k / (double)n0;
        }
        int64_t e_cap = (int64_t)src.size() + (n_final - n0) * (int64_t)std::max(m, 1) + 16;
        if (model == "er-gnp")
            e_cap = (int64_t)src.size() + (int64_t)(1.2 * p_er * (double)n_final * (double)(n_final - n0)) + 1000000;
        if (ref_model)  // at most max(k, 2 + floor(k)) edges per paper (generator and same-year included)
            e_cap = (int64_t)src.size() + (n_final - n0) * (int64_t)(max_out + 2) + 16;
        std::vector<float> logfit(n_final);
        if (pg_seed_logfit(seed, 0, n_final, 0.0, a.num("--fitness-sigma", 0.5), logfit.data()))
            throw std::runtime_error("fitness draw failed");
        pg_config cfg{};
        cfg.size = sizeof(cfg);
        cfg.device = (int)a.num("--device", 0);
        cfg.max_papers = n_final;
        cfg.max_edges = e_cap;
        cfg.max_new_per_year = std::max<int64_t>(max_new, 1);
        cfg.max_cites = std::max(m, 1);
        cfg.bfs_cap = (int)a.num("--bfs-cap", 4096);
        cfg.seed = seed;
        pg_ctx* ctx = nullptr;
        check(pg_ctx_create(&cfg, &ctx), nullptr, "pg_ctx_create");
        check(pg_graph_load(ctx, n0, birth.data(), logfit.data(), n_final, (int64_t)src.size(),
                            src.data(), dst.data()), ctx, "pg_graph_load");
        log.info("loaded " + std::to_string(n0) + " papers and " + std::to_string(src.size()) +
                 " citations");
        std::vector<int32_t> roots(n_final - n0, -1);
        int y_start = 0;
        if (a.has("--resume-from")) {
            // the checkpoint replaces the loaded state; its blob is the roots so far
            int64_t ub = 0;
            check(pg_checkpoint_load(ctx, a.str("--resume-from").c_str(), roots.data(),
                                     (int64_t)roots.size() * 4, &ub), ctx, "pg_checkpoint_load");
            int32_t done = 0, last_year = 0, last_yi = -1;
            pg_get_progress(ctx, &done, &last_year, &last_yi);
            int64_t grown = 0;
            for (int y = 0; y < std::min(done, years); ++y) grown += sched[y];
            if (done > years || last_yi != done - 1 || (done > 0 && last_year != max_year + done) ||
                ub != grown * 4)
                throw std::runtime_error("--resume-from: the checkpoint is not of this run");
            y_start = done;
            log.info("resumed after " + std::to_string(done) + " years");
        }
        if (ref_model) {
            pg_abm_model am{};
            am.size = sizeof(am);
            am.bag = bag.data();
            am.n_bag = (int32_t)bag.size();
            am.rec_prob = rec.empty() ? nullptr : rec.data();
            am.n_rec = (int32_t)rec.size();
            am.alpha = a.num("--alpha", -1.0);
            am.w_pa = a.num("--preferential-weight", -1.0);
            am.w_rec = a.num("--recency-weight", -1.0);
            am.w_fit = a.num("--fitness-weight", -1.0);
            am.fully_random = a.num("--fully-random-citations", 0.05);
            am.same_year = a.num("--same-year-proportion", 0.0);
            check(pg_abm_configure(ctx, &am), ctx, "pg_abm_configure");
        }
        // ---- year loop ----
        const auto t_loop = Clock::now();
        int64_t total_edges = 0;
        for (int y = y_start; y < years; ++y) {
            pg_year_params p{};
            p.size = sizeof(p);
            p.model = ref_model ? PG_MODEL_ABM : model == "pa" ? PG_MODEL_PA : PG_MODEL_ER;
            p.year = max_year + 1 + y;
            p.year_index = y;
            p.n_new = sched[y];
            p.m = m;
            p.n_nb = n_nb;
            p.alpha = (float)alpha;
            p.inv_tau = std::isinf(tau) ? 0.0f : (float)(1.0 / tau);
            p.er_p = p_er;
            pg_year_stats st{};
            st.size = sizeof(st);
            int64_t nb_before = 0, e_unused = 0;
            pg_graph_size(ctx, &nb_before, &e_unused);
            check(pg_year_step(ctx, &p, &st), ctx, "pg_year_step");
            if (ref_model && sched[y] > 0) {
                std::vector<int32_t> info((size_t)sched[y] * 6);
                check(pg_abm_year_info(ctx, info.data()), ctx, "pg_abm_year_info");
                for (int64_t i = 0; i < sched[y]; ++i) roots[nb_before - n0 + i] = info[(size_t)i * 6];
            }
            if (model == "pa" && m > 0) {
                std::vector<int32_t> cit((size_t)sched[y] * m);
                check(pg_year_citations(ctx, cit.data(), nullptr, nullptr, nullptr), ctx,
                      "pg_year_citations");
                for (int64_t i = 0; i < sched[y]; ++i) roots[nb_before - n0 + i] = cit[(size_t)i * m];
            }
            total_edges += st.edges_added;
            char msg[256];
            std::snprintf(msg, sizeof(msg),
                          "Year %d: N=%lld new=%lld edges=%lld near_boundary=%lld shortfall=%lld "
                          "device_ms=%.2f", p.year, (long long)st.n_before, (long long)sched[y],
                          (long long)st.edges_added, (long long)st.near_boundary,
                          (long long)st.shortfall, st.device_ms);
            log.info(msg);
            std::printf("%s\n", msg);
            if (a.has("--checkpoint-file"))  // roots of the papers grown so far as the blob
                check(pg_checkpoint_save(ctx, a.str("--checkpoint-file").c_str(), roots.data(),
                                         (nb_before + sched[y] - n0) * 4), ctx, "pg_checkpoint_save");
        }
        const double loop_s = std::chrono::duration<double>(Clock::now() - t_loop).count();
        std::printf("Growth: %d years, %lld new citation edges in %.3f s (%.3e edges/s)\n", years - y_start,
                    (long long)total_edges, loop_s, loop_s > 0 ? total_edges / loop_s : 0.0);
        // ---- outputs ----
        int64_t nn = 0, ne = 0;
        pg_graph_size(ctx, &nn, &ne);
        std::vector<int32_t> fo(nn + 1), fa(ne), bo(nn + 1), indeg(nn), by(nn);
        check(pg_graph_download(ctx, fo.data(), fa.data(), bo.data(), nullptr, indeg.data(),
                                by.data(), nullptr), ctx, "pg_graph_download");
        // the scoring model's static draws of every paper (aux columns)
        std::vector<int32_t> r_k, r_fit;
        std::vector<double> r_alpha, r_w;
        if (ref_model) {
            r_k.resize(nn);
            r_fit.resize(nn);
            r_alpha.resize(nn);
            r_w.resize((size_t)nn * 3);
            parallel_for(T, [&](int t) {
                const int64_t b = nn / T * t, e = t + 1 == T ? nn : nn / T * (t + 1);
                check(pg_abm_paper_draws(ctx, b, e - b, r_k.data() + b, r_alpha.data() + b,
                                         r_w.data() + 3 * b, r_fit.data() + b), ctx,
                      "pg_abm_paper_draws");
            });
        }
        pg_ctx_destroy(ctx);
        long long next_id = orig[n0 - 1] + 1;
        orig.resize(nn);
        for (int64_t v = n0; v < nn; ++v) orig[v] = next_id++;
        const auto t_write = Clock::now();
        {
            // rows of papers [pb[t], pb[t+1]) formatted by thread t, about
            // ne / T edges each
            std::vector<int64_t> pb(T + 1, nn);
            pb[0] = 0;
            for (int t = 1; t < T; ++t) {
                const int64_t target = ne / T * t;
                pb[t] = std::max<int64_t>(pb[t - 1],
                                          std::upper_bound(fo.begin(), fo.begin() + nn, target) -
                                              fo.begin() - 1);
            }
            std::vector<std::string> part(T);
            parallel_for(T, [&](int t) {
                std::string& buf = part[t];
                buf.reserve((size_t)(fo[pb[t + 1]] - fo[pb[t]]) * 16 + 16);
                for (int64_t v = pb[t]; v < pb[t + 1]; ++v) {
                    for (int32_t k = fo[v]; k < fo[v + 1]; ++k) {
                        append_int(buf, orig[v]);
                        buf.push_back(',');
                        append_int(buf, orig[fa[k]]);
                        buf.push_back('\n');
                    }
                }
            });
            std::ofstream out(a.str("--output-file"), std::ios::binary);
            out << "#source,target\n";
            for (const auto& b : part) out.write(b.data(), (std::streamsize)b.size());
            if (!out) throw std::runtime_error("cannot write --output-file");
        }
        {
            std::vector<std::string> part(T);
            parallel_for(T, [&](int t) {
                const int64_t b = nn / T * t, e = t + 1 == T ? nn : nn / T * (t + 1);
                std::string& buf = part[t];
                buf.reserve((size_t)(e - b) * 64);
                char row[320];
                for (int64_t v = b; v < e; ++v) {
                    const int outd = fo[v + 1] - fo[v];
                    int len;
                    if (ref_model && v < n0) {
                        // the reference's seed rows: static fitness (lag 0,
                        // peak duration 1000, src/abm.cu:124-150)
                        len = std::snprintf(row, sizeof(row), "%lld,seed,%d,-1,-1,-1,0,%d,1000,-1,%d,%d,-1,-1,no_generators\n",
                                            orig[v], by[v], r_fit[v], indeg[v], outd);
                    } else if (ref_model) {
                        const int32_t r = roots[v - n0];
                        const long long ro = r >= 0 ? orig[r] : -1;
                        len = std::snprintf(row, sizeof(row), "%lld,agent,%d,%.10g,%.10g,%.10g,0,%d,1000,%.10g,%d,%d,%d,-1,%lld\n",
                                            orig[v], by[v], r_w[3 * v], r_w[3 * v + 1], r_w[3 * v + 2],
                                            r_fit[v], r_alpha[v], indeg[v], outd, r_k[v], ro);
                    } else if (v < n0) {
                        len = std::snprintf(row, sizeof(row), "%lld,seed,%d,-1,-1,-1,0,0,0,-1,%d,%d,-1,-1,no_generators\n",
                                            orig[v], by[v], indeg[v], outd);
                    } else {
                        const int32_t r = roots[v - n0];
                        const long long ro = r >= 0 ? orig[r] : -1;
                        len = std::snprintf(row, sizeof(row), "%lld,agent,%d,%.10g,%.10g,%.10g,0,0,0,%.10g,%d,%d,%d,-1,%lld\n",
                                            orig[v], by[v], alpha, tau, std::exp((double)logfit[v]),
                                            f_nb, indeg[v], outd, m, ro);
                    }
                    buf.append(row, (size_t)len);
                }
            });
            std::ofstream out(a.str("--auxiliary-information-file"), std::ios::binary);
            out << "node_id,type,year,pa_weight,rec_weight,fit_weight,fit_lag_duration,"
                   "fit_peak_value,fit_peak_duration,alpha,in_degree,out_degree,"
                   "assigned_out_degree,planted_nodes_line_number,generator_node_string\n";
            for (const auto& b : part) out.write(b.data(), (std::streamsize)b.size());
            if (!out) throw std::runtime_error("cannot write --auxiliary-information-file");
        }
        if (a.has("--output-nodelist")) {
            // `node,year` per paper in ascending original id (the order the
            // nodelist parser assigns dense ids in, src/graph.cu:336-373), so
            // this file with --output-file is the input of a chained run
            std::vector<std::string> part(T);
            parallel_for(T, [&](int t) {
                const int64_t b = nn / T * t, e = t + 1 == T ? nn : nn / T * (t + 1);
                std::string& buf = part[t];
                buf.reserve((size_t)(e - b) * 16);
                for (int64_t v = b; v < e; ++v) {
                    append_int(buf, orig[v]);
                    buf.push_back(',');
                    append_int(buf, by[v]);
                    buf.push_back('\n');
                }
            });
            std::ofstream out(a.str("--output-nodelist"), std::ios::binary);
            out << "#node,year\n";
            for (const auto& b : part) out.write(b.data(), (std::streamsize)b.size());
            if (!out) throw std::runtime_error("cannot write --output-nodelist");
        }
        const double write_s = std::chrono::duration<double>(Clock::now() - t_write).count();
        std::printf("I/O: read %.3f s, wrote %.3f s (%d threads)\n", read_s, write_s, T);
        log.info("wrote " + std::to_string(ne) + " edges for " + std::to_string(nn) + " papers");
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "abm: %s\n", e.what());
        return 1;
    }
}
