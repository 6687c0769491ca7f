// Tecplot BLOCK writer (block_writer.h): Fortran ES23.16 formatting, the
// record layout of the reference's output formats, and the writer thread.
#include "block_writer.h"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <future>

namespace hugks {

void format_es23_16(double v, char* out) {
    memset(out, ' ', 23);
    if (std::isnan(v) || std::isinf(v)) {
        const char* s = std::isnan(v) ? "NaN" : (v < 0 ? "-Inf" : "Inf");
        const size_t n = strlen(s);
        memcpy(out + 23 - n, s, n);
        return;
    }
    // "%.16E" is exact decimal conversion, correctly rounded: -d.dddd...E+dd
    char buf[40];
    const int n = snprintf(buf, sizeof buf, "%.16E", v);
    const char* e = strchr(buf, 'E');
    const int mant = (int)(e - buf);           // sign (if any) and d.dddd
    const int edig = n - mant - 2;             // exponent digits after E and its sign
    char tail[6];
    int tl;
    if (edig <= 2) {
        tail[0] = 'E';
        tail[1] = e[1];
        tail[2] = edig == 2 ? e[2] : '0';
        tail[3] = edig == 2 ? e[3] : e[2];
        tl = 4;
    } else {
        // |exponent| > 99: the exponent letter is dropped (ESw.d without Ee)
        tail[0] = e[1];
        memcpy(tail + 1, e + 2, (size_t)edig);
        tl = 1 + edig;
    }
    const int len = mant + tl;
    char* p = out + 23 - len;
    memcpy(p, buf, (size_t)mant);
    memcpy(p + mant, tail, (size_t)tl);
}

// the records of values[0..n) into s, per_line fields per record
static void format_records(const double* v, int64_t n, int per_line, std::string& s) {
    s.clear();
    s.reserve((size_t)n * 25 + (size_t)(n / per_line + 1));
    char f[23];
    for (int64_t k = 0; k < n; ++k) {
        const int col = (int)(k % per_line);
        if (col > 0) s.append("  ", 2);
        format_es23_16(v[k], f);
        s.append(f, 23);
        if (col == per_line - 1 || k == n - 1) s.push_back('\n');
    }
}

std::string write_block_file(const BlockFile& f) {
    // each block formatted in chunks of whole records (in parallel, at most
    // eight chunks at a time), then written in order; an empty block is one
    // empty record, as a Fortran write of an empty array makes
    constexpr int64_t kChunkRecords = 16384;
    const int64_t per_chunk = kChunkRecords * f.per_line;
    struct Chunk {
        const double* src;
        int64_t n;
    };
    std::vector<Chunk> chunks;
    int64_t off = 0;
    for (int64_t len : f.block_len) {
        if (len == 0) chunks.push_back({nullptr, 0});
        for (int64_t a = 0; a < len; a += per_chunk)
            chunks.push_back({f.values.data() + off + a, len - a < per_chunk ? len - a : per_chunk});
        off += len;
    }
    std::vector<std::string> parts(chunks.size());
    std::vector<std::future<void>> jobs;
    for (size_t i = 0; i < chunks.size(); ++i) {
        if (chunks[i].n == 0) {
            parts[i] = "\n";
            continue;
        }
        const Chunk ch = chunks[i];
        std::string* dst = &parts[i];
        const int pl = f.per_line;
        jobs.push_back(std::async(std::launch::async,
                                  [ch, pl, dst] { format_records(ch.src, ch.n, pl, *dst); }));
        if (jobs.size() >= 8) {
            jobs.front().get();
            jobs.erase(jobs.begin());
        }
    }
    for (auto& j : jobs) j.get();
    FILE* fp = fopen(f.path.c_str(), "w");
    if (!fp) return f.path + ": " + strerror(errno);
    bool ok = fwrite(f.header.data(), 1, f.header.size(), fp) == f.header.size();
    for (const std::string& p : parts) ok = ok && fwrite(p.data(), 1, p.size(), fp) == p.size();
    ok = (fclose(fp) == 0) && ok;
    return ok ? std::string() : f.path + ": write failed";
}

BlockWriter::~BlockWriter() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    if (th_.joinable()) th_.join();
}

void BlockWriter::push(BlockFile&& f) {
    {
        std::lock_guard<std::mutex> lk(mu_);
        q_.push_back(std::move(f));
        if (!th_.joinable()) th_ = std::thread(&BlockWriter::run, this);
    }
    cv_.notify_all();
}

std::string BlockWriter::wait() {
    std::unique_lock<std::mutex> lk(mu_);
    idle_cv_.wait(lk, [this] { return q_.empty() && !busy_; });
    std::string e;
    e.swap(err_);
    return e;
}

void BlockWriter::run() {
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        cv_.wait(lk, [this] { return stop_ || !q_.empty(); });
        if (q_.empty()) break;   // stop requested and nothing left
        BlockFile f = std::move(q_.front());
        q_.pop_front();
        busy_ = true;
        lk.unlock();
        const std::string e = write_block_file(f);
        lk.lock();
        busy_ = false;
        if (!e.empty() && err_.empty()) err_ = e;
        if (q_.empty()) idle_cv_.notify_all();
    }
    idle_cv_.notify_all();
}

}  // namespace hugks
