// Tecplot BLOCK output off the step's critical path: a queue of files and
// one writer thread per owner (a solver handle).  The caller's values are
// copied when a file is queued, so the caller may reuse its buffers and go
// on stepping while the file is formatted and written.
//
// The records are those the reference's writers make with the Fortran
// formats "(11(ES23.16,2X))" (2d/src/output2d.f90:48-53) and
// "(9(ES23.16,2X))" (1d/src/output.f90:95-97): per_line fields per record,
// each ES23.16 (see format_es23_16), two blanks between fields and none
// after the last, every block starting a new record.
#pragma once
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace hugks {

// One value in Fortran ES23.16 form into out[0..22]: the sign or a blank,
// d.dddddddddddddddd (correctly rounded), then E+dd / E-dd, or +ddd / -ddd
// without the E for decimal exponents beyond 99 (exact decimal conversion,
// ties to even); NaN, Inf and -Inf right-justified -- as the Fortran
// runtime of the reference build writes them.
void format_es23_16(double v, char* out);

struct BlockFile {
    std::string path, header;
    int per_line = 0;
    std::vector<int64_t> block_len;
    std::vector<double> values;
};

// Formats and writes one file; returns an empty string or the error.
std::string write_block_file(const BlockFile& f);

class BlockWriter {
public:
    BlockWriter() = default;
    BlockWriter(const BlockWriter&) = delete;
    BlockWriter& operator=(const BlockWriter&) = delete;
    ~BlockWriter();
    void push(BlockFile&& f);
    // waits until the queue is empty and nothing is being written; returns
    // the first error since the last wait (empty when none)
    std::string wait();

private:
    void run();
    std::mutex mu_;
    std::condition_variable cv_, idle_cv_;
    std::deque<BlockFile> q_;
    std::thread th_;
    bool busy_ = false, stop_ = false;
    std::string err_;
};

}  // namespace hugks
