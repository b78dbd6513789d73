// katoann: the command-line front end of the gfx950 engine, over libkatoann's
// C ABI (through the C++ API of include/katoann/index.h).
//
// Commands, flags, files and the JSON line each command prints are the
// reference's (tools/katoann.cpp: build :201-219, search :221-291, load
// :296-339, usage :341-355), so scripts written against it run unchanged:
//   build   --data BASE.fbin --out PREFIX [--nlist --train-size --iters --seed
//           --normalize | --normalized --threads]
//   search  --index PREFIX --queries Q.fbin [--gt --k --nprobe a,b,.. --ranking
//           --survivors --prefix-dims --selection --batch --precision --engine
//           --coarse --repeat --threads --query-ids --out]
//   load    --index PREFIX --queries Q.fbin --rate QPS [--requests --nprobe
//           --max-batch --seed] and the search options
// Differences: the engine is the GPU, so --engine takes amx, avx512 or gfx950
// and reports "gfx950", and --threads is accepted and ignored ("threads"
// reports the one host thread that drives the GPU). load --inflight N (N > 1,
// not in the reference) serves the same arrivals through the continuous-
// batching server with N batches on the GPU at once; its line adds "inflight".
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <random>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "katoann/index.h"

namespace {

using Clock = std::chrono::steady_clock;

double since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

// --name value, --name=value, or a bare --normalize / --normalized.
class Options {
 public:
  Options(int argc, char **argv, int first) {
    for (int i = first; i < argc; ++i) {
      const std::string a = argv[i];
      if (a.size() < 3 || a.compare(0, 2, "--") != 0) throw std::invalid_argument("unexpected argument " + a);
      const std::string body = a.substr(2);
      const size_t eq = body.find('=');
      if (eq != std::string::npos) {
        values_[body.substr(0, eq)] = body.substr(eq + 1);
      } else if (body == "normalize" || body == "normalized") {
        switches_.insert(body);
      } else if (i + 1 < argc) {
        values_[body] = argv[++i];
      } else {
        throw std::invalid_argument("missing value for --" + body);
      }
    }
  }
  bool has(const std::string &k) const { return values_.count(k) != 0; }
  bool on(const std::string &k) const { return switches_.count(k) != 0; }
  std::string str(const std::string &k, const std::string &dflt = "") const {
    const auto it = values_.find(k);
    return it == values_.end() ? dflt : it->second;
  }
  std::string need(const std::string &k) const {
    if (!has(k)) throw std::invalid_argument("missing --" + k);
    return values_.at(k);
  }
  uint32_t u32(const std::string &k, uint32_t dflt) const {
    return has(k) ? static_cast<uint32_t>(std::stoul(values_.at(k))) : dflt;
  }
  double f64(const std::string &k, double dflt) const { return has(k) ? std::stod(values_.at(k)) : dflt; }

 private:
  std::map<std::string, std::string> values_;
  std::set<std::string> switches_;
};

std::vector<uint32_t> numbers(const std::string &csv) {
  std::vector<uint32_t> out;
  std::stringstream in(csv);
  for (std::string tok; std::getline(in, tok, ',');) out.push_back(static_cast<uint32_t>(std::stoul(tok)));
  return out;
}

// u32 rows, u32 cols, then rows x cols values of T (.fbin vectors; ground truth ids).
template <class T>
std::vector<T> read_table(const std::string &path, uint32_t &rows, uint32_t &cols) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  uint32_t hdr[2] = {0, 0};
  f.read(reinterpret_cast<char *>(hdr), sizeof(hdr));
  rows = hdr[0];
  cols = hdr[1];
  std::vector<T> v(static_cast<size_t>(rows) * cols);
  f.read(reinterpret_cast<char *>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(T)));
  if (!f) throw std::runtime_error("truncated " + path);
  return v;
}

// Unit rows (the front ends' rule: fp64 sum of squares, fp32 1/sqrt, fp32
// products), computed by the library on the GPU.
void unit_rows(std::vector<float> &v, uint32_t n, uint32_t d) {
  if (n) katoann::detail::check(katoann_normalize_rows(v.data(), n, d));
}

katoann::LoadOptions load_options(const Options &o) {
  katoann::LoadOptions lo;
  const std::string prec = o.str("precision", "sq16");
  if (prec == "bf16") lo.precision = katoann::Precision::kBf16;
  else if (prec != "sq16") throw std::invalid_argument("--precision is sq16 or bf16");
  const std::string eng = o.str("engine", "amx");
  if (eng != "amx" && eng != "avx512" && eng != "gfx950")
    throw std::invalid_argument("--engine is amx, avx512 or gfx950");
  lo.prefix_dims = o.u32("prefix-dims", 0);
  const std::string coarse = o.str("coarse", "batched");
  if (coarse == "flat") lo.batched_coarse = false;
  else if (coarse != "batched") throw std::invalid_argument("--coarse is batched or flat");
  return lo;
}

katoann::SearchParams search_params(const Options &o, const katoann::Index &ix) {
  katoann::SearchParams p;
  p.k = o.u32("k", 10);
  const std::string r = o.str("ranking", ix.precision() == katoann::Precision::kBf16 ? "full" : "prefix");
  if (r == "prefix") p.ranking = katoann::Ranking::kPrefix;
  else if (r == "high") p.ranking = katoann::Ranking::kHighBytes;
  else if (r == "full") p.ranking = katoann::Ranking::kFull;
  else throw std::invalid_argument("--ranking is prefix, high, or full");
  p.survivors = o.u32("survivors", 32);
  const std::string s = o.str("selection", "threshold");
  if (s == "local") p.selection = katoann::Selection::kLocal;
  else if (s != "threshold") throw std::invalid_argument("--selection is threshold or local");
  return p;
}

// The queries to run: every row, or the ids listed in --query-ids.
std::vector<uint32_t> chosen_queries(const Options &o, uint32_t nq) {
  std::vector<uint32_t> ids;
  if (!o.has("query-ids")) {
    ids.resize(nq);
    for (uint32_t i = 0; i < nq; ++i) ids[i] = i;
    return ids;
  }
  const std::string path = o.str("query-ids");
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open " + path);
  for (uint32_t v; f >> v;) {
    if (v >= nq) throw std::runtime_error("query id out of range in " + path);
    ids.push_back(v);
  }
  if (ids.empty()) throw std::runtime_error("no query ids in " + path);
  return ids;
}

// Rows `which` of the table, contiguous.
std::vector<float> take_rows(const std::vector<float> &all, uint32_t d, const std::vector<uint32_t> &which) {
  std::vector<float> out(which.size() * static_cast<size_t>(d));
  for (size_t i = 0; i < which.size(); ++i)
    std::copy_n(all.begin() + static_cast<size_t>(which[i]) * d, d, out.begin() + i * d);
  return out;
}

double recall(const std::vector<uint32_t> &ids, uint32_t k, const std::vector<uint32_t> &which,
              const std::vector<uint32_t> &gt, uint32_t gk) {
  uint64_t hits = 0;
  for (size_t i = 0; i < which.size(); ++i) {
    const uint32_t *truth = gt.data() + static_cast<size_t>(which[i]) * gk;
    for (uint32_t j = 0; j < k; ++j)
      hits += std::find(truth, truth + k, ids[i * k + j]) != truth + k ? 1 : 0;
  }
  return static_cast<double>(hits) / (static_cast<double>(which.size()) * k);
}

// The q-quantile of sorted values: sorted[min(n - 1, ceil(q n) - 1)].
double quantile(const std::vector<double> &sorted, double q) {
  const size_t n = sorted.size();
  const double r = std::ceil(q * static_cast<double>(n));
  const size_t i = r < 1.0 ? 0 : static_cast<size_t>(r) - 1;
  return sorted[std::min(n - 1, i)];
}

int run_build(const Options &o) {
  uint32_t n = 0, d = 0;
  std::vector<float> data = read_table<float>(o.need("data"), n, d);
  katoann::BuildOptions bo;
  bo.nlist = o.u32("nlist", 2048);
  bo.train_size = o.u32("train-size", 256000);
  bo.kmeans_iters = o.u32("iters", 12);
  bo.seed = std::stoull(o.str("seed", "0"));
  if (o.on("normalize")) unit_rows(data, n, d);
  bo.normalized = o.on("normalize") || o.on("normalized");
  const std::string out = o.need("out");
  const auto t0 = Clock::now();
  katoann::Index::build(data.data(), n, d, bo, out);
  std::printf("{\"vectors\": %u, \"dim\": %u, \"nlist\": %u, \"build_s\": %.1f}\n", n, d, bo.nlist, since(t0));
  return 0;
}

struct Loaded {
  katoann::Index ix;
  double load_s = 0;
  uint32_t nq = 0, d = 0;
  std::vector<float> queries;  // all rows, unit length if the index is
};

void load_index_and_queries(const Options &o, Loaded &L) {
  const auto t0 = Clock::now();
  L.ix.load(o.need("index"), load_options(o));
  L.load_s = since(t0);
  L.queries = read_table<float>(o.need("queries"), L.nq, L.d);
  if (L.d != L.ix.dim()) throw std::runtime_error("query dimension does not match the index");
  if (L.ix.normalized()) unit_rows(L.queries, L.nq, L.d);
}

int run_search(const Options &o) {
  Loaded L;
  load_index_and_queries(o, L);
  const std::vector<uint32_t> which = chosen_queries(o, L.nq);
  const std::vector<float> q = take_rows(L.queries, L.d, which);
  uint32_t gn = 0, gk = 0;
  std::vector<uint32_t> gt;
  if (!o.str("gt").empty()) gt = read_table<uint32_t>(o.str("gt"), gn, gk);
  katoann::SearchParams p = search_params(o, L.ix);
  if (!gt.empty() && (gn < L.nq || gk < p.k)) throw std::runtime_error("ground truth too small");
  const uint32_t batch = std::max(1u, o.u32("batch", 4096));
  const uint32_t repeat = o.u32("repeat", 3);
  const std::string out = o.str("out");
  std::printf("{\"index\": \"%s\", \"engine\": \"gfx950\", \"threads\": 1, \"load_s\": %.1f, "
              "\"memory_gb\": %.2f, \"queries\": %zu, \"prefix_dims\": %u}\n",
              o.need("index").c_str(), L.load_s, static_cast<double>(L.ix.memory_bytes()) / 1e9, which.size(),
              L.ix.prefix_dims());
  std::fflush(stdout);
  const size_t total = which.size();
  std::vector<uint32_t> ids(total * p.k);
  std::vector<float> dists(total * p.k);
  for (const uint32_t np : numbers(o.str("nprobe", "64"))) {
    p.nprobe = np;
    auto one_pass = [&] {
      for (size_t at = 0; at < total; at += batch) {
        const uint32_t b = static_cast<uint32_t>(std::min<size_t>(batch, total - at));
        L.ix.search(q.data() + at * L.d, b, p, ids.data() + at * p.k, dists.data() + at * p.k);
      }
    };
    one_pass();  // untimed
    std::vector<double> secs;
    for (uint32_t r = 0; r < repeat; ++r) {
      const auto t0 = Clock::now();
      one_pass();
      secs.push_back(since(t0));
    }
    if (secs.empty()) secs.push_back(0.0);
    std::sort(secs.begin(), secs.end());
    const double med = secs[secs.size() / 2];
    std::printf("{\"nprobe\": %u, \"k\": %u, \"survivors\": %u, \"batch\": %u, \"qps\": %.1f, "
                "\"seconds_median\": %.4f, \"seconds_min\": %.4f, \"seconds_max\": %.4f",
                np, p.k, p.survivors, batch, med > 0 ? static_cast<double>(total) / med : 0.0, med, secs.front(),
                secs.back());
    if (!gt.empty()) std::printf(", \"recall\": %.5f", recall(ids, p.k, which, gt, gk));
    std::printf("}\n");
    std::fflush(stdout);
    if (!out.empty()) {  // u32 queries, u32 k, ids, then distances
      const std::string path = out + ".np" + std::to_string(np) + ".bin";
      std::ofstream f(path, std::ios::binary);
      const uint32_t hdr[2] = {static_cast<uint32_t>(total), p.k};
      f.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
      f.write(reinterpret_cast<const char *>(ids.data()), static_cast<std::streamsize>(ids.size() * 4));
      f.write(reinterpret_cast<const char *>(dists.data()), static_cast<std::streamsize>(dists.size() * 4));
      if (!f) throw std::runtime_error("cannot write " + path);
    }
  }
  return 0;
}

// Open-loop serving: request i (one query, row pool[i % |pool|]) arrives at the
// i-th point of a Poisson process of the given rate. With one batch in flight
// (the reference's policy), whenever the GPU is idle every request that has
// arrived, up to --max-batch, is searched as one batch; latency runs from
// arrival to the end of that batch.
int run_load(const Options &o) {
  Loaded L;
  load_index_and_queries(o, L);
  const std::vector<uint32_t> pool = chosen_queries(o, L.nq);
  katoann::SearchParams p = search_params(o, L.ix);
  p.nprobe = o.u32("nprobe", 64);
  const double rate = o.f64("rate", 1000);
  const uint32_t n = o.u32("requests", 10000);
  const uint32_t max_batch = std::max(1u, o.u32("max-batch", 4096));
  const uint32_t inflight = std::max(1u, o.u32("inflight", 1));
  if (n == 0) throw std::invalid_argument("--requests must be positive");
  std::mt19937_64 gen(o.u32("seed", 42));
  std::exponential_distribution<double> gap(rate);
  std::vector<double> arrive(n);
  double t = 0;
  for (uint32_t i = 0; i < n; ++i) arrive[i] = (t += gap(gen));
  const uint32_t d = L.d;
  auto row = [&](uint32_t i) { return L.queries.data() + static_cast<size_t>(pool[i % pool.size()]) * d; };
  std::vector<float> bq(static_cast<size_t>(max_batch) * d);
  std::vector<uint32_t> ids(static_cast<size_t>(max_batch) * p.k);
  std::vector<float> dists(static_cast<size_t>(max_batch) * p.k);
  {  // untimed: one full batch
    const uint32_t w = std::min<uint32_t>(max_batch, static_cast<uint32_t>(pool.size()));
    for (uint32_t i = 0; i < w; ++i) std::copy_n(row(i), d, bq.begin() + static_cast<size_t>(i) * d);
    L.ix.search(bq.data(), w, p, ids.data(), dists.data());
  }
  std::vector<double> lat(n);
  size_t batches = 0;
  double mean_batch = 0;
  const auto start = Clock::now();
  if (inflight == 1) {
    for (uint32_t next = 0; next < n;) {
      double now = since(start);
      if (arrive[next] > now) {
        std::this_thread::sleep_for(std::chrono::duration<double>(arrive[next] - now));
        now = since(start);
      }
      uint32_t end = next;
      while (end < n && end - next < max_batch && arrive[end] <= now) ++end;
      const uint32_t b = end - next;
      for (uint32_t i = 0; i < b; ++i) std::copy_n(row(next + i), d, bq.begin() + static_cast<size_t>(i) * d);
      L.ix.search(bq.data(), b, p, ids.data(), dists.data());
      const double done = since(start);
      for (uint32_t i = next; i < end; ++i) lat[i] = done - arrive[i];
      mean_batch += b;
      ++batches;
      next = end;
    }
    mean_batch /= static_cast<double>(batches);
  } else {
    // continuous batching: each request is submitted at its arrival; the
    // server forms batches of what is queued and keeps `inflight` on the GPU
    katoann_server_options so;
    katoann_default_server_options(&so);
    so.max_inflight = inflight;
    so.max_batch = max_batch;
    katoann_server *srv = nullptr;
    katoann::detail::check(katoann_server_create(L.ix.handle(), &so, &srv));
    const katoann_search_params sp{p.k, p.nprobe,
                                   p.ranking == katoann::Ranking::kFull        ? KATOANN_RANK_FULL
                                   : p.ranking == katoann::Ranking::kHighBytes ? KATOANN_RANK_HIGH
                                                                               : KATOANN_RANK_PREFIX,
                                   p.survivors,
                                   p.selection == katoann::Selection::kLocal ? KATOANN_SEL_LOCAL
                                                                             : KATOANN_SEL_THRESHOLD};
    std::vector<uint64_t> ticket(n);
    std::vector<double> submitted(n);
    std::atomic<uint32_t> posted{0};
    std::string err, err_submit;
    std::thread producer([&] {
      for (uint32_t i = 0; i < n; ++i) {
        const double now = since(start);
        if (arrive[i] > now) std::this_thread::sleep_for(std::chrono::duration<double>(arrive[i] - now));
        submitted[i] = since(start);
        if (katoann_server_submit(srv, row(i), 1, &sp, nullptr, 0, &ticket[i]) != KATOANN_OK) {
          err_submit = katoann_last_error();
          break;
        }
        posted.store(i + 1, std::memory_order_release);
      }
      posted.store(n + 1, std::memory_order_release);  // done (or failed)
    });
    std::vector<uint32_t> one_id(p.k);
    std::vector<float> one_d(p.k);
    for (uint32_t i = 0; i < n; ++i) {
      while (posted.load(std::memory_order_acquire) <= i) std::this_thread::yield();
      if (!err_submit.empty()) break;
      double ms = 0;
      const int rc = katoann_server_wait(srv, ticket[i], one_id.data(), one_d.data(), &ms);
      if (rc != KATOANN_OK) {
        err = katoann_last_error();
        break;
      }
      // the server times from submission; add the lag behind the arrival
      lat[i] = submitted[i] - arrive[i] + ms / 1e3;
    }
    producer.join();
    if (err.empty()) err = err_submit;
    katoann_server_stats st{};
    katoann_server_stats_get(srv, &st);
    katoann_server_destroy(srv);
    if (!err.empty()) throw std::runtime_error(err);
    batches = st.batches;
    mean_batch = st.mean_batch;
  }
  // span: first arrival-clock start to the last completion
  double span = 0;
  for (uint32_t i = 0; i < n; ++i) span = std::max(span, arrive[i] + lat[i]);
  std::vector<double> sorted = lat;
  std::sort(sorted.begin(), sorted.end());
  std::printf("{\"rate\": %.1f, \"requests\": %u, \"nprobe\": %u, \"engine\": \"gfx950\", "
              "\"achieved_qps\": %.1f, \"p50_ms\": %.2f, \"p95_ms\": %.2f, \"p99_ms\": %.2f, "
              "\"max_ms\": %.2f, \"batches\": %zu, \"mean_batch\": %.1f",
              rate, n, p.nprobe, n / span, quantile(sorted, 0.50) * 1e3, quantile(sorted, 0.95) * 1e3,
              quantile(sorted, 0.99) * 1e3, sorted.back() * 1e3, batches, mean_batch);
  if (inflight > 1) std::printf(", \"inflight\": %u", inflight);
  std::printf("}\n");
  return 0;
}

void usage() {
  std::fputs(
      "usage: katoann build  --data BASE.fbin --out PREFIX [--nlist 2048] [--normalize]\n"
      "                      [--train-size 256000] [--iters 12] [--seed 0] [--threads T]\n"
      "       katoann search --index PREFIX --queries Q.fbin [--gt GT.bin] [--k 10]\n"
      "                      [--nprobe 64,128] [--ranking prefix|high|full] [--survivors 32]\n"
      "                      [--prefix-dims D] [--selection threshold|local] [--batch 4096]\n"
      "                      [--precision sq16|bf16] [--engine amx|avx512|gfx950]\n"
      "                      [--coarse batched|flat] [--repeat 3] [--threads T]\n"
      "                      [--query-ids FILE] [--out PREFIX]\n"
      "       katoann load   --index PREFIX --queries Q.fbin --rate QPS [--requests 10000]\n"
      "                      [--nprobe 64] [--max-batch 4096] [--seed 42] [--inflight 1]\n"
      "                      and the search options above\n",
      stderr);
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    usage();
    return 2;
  }
  const std::string cmd = argv[1];
  try {
    const Options o(argc, argv, 2);
    if (cmd == "build") return run_build(o);
    if (cmd == "search") return run_search(o);
    if (cmd == "load") return run_load(o);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "katoann: %s\n", e.what());
    return 1;
  }
  usage();
  return 2;
}
