// lcs_io.h -- the reference's on-disk formats, written from host copies of
// results kept in HBM, by a background thread so that the GPU keeps working.
//
//   * FTLE output (write_lcs / write_ftle / write_data_from_master,
//     io_nohdf5_m.f90:47-332): a Tecplot ASCII file, DATAPACKING = BLOCK,
//     ES14.6 values, 10 per line, every variable starting on a new line.
//   * Flowmap / velocity slot files (write_flowmap_substep,
//     flowmap_m.f90:106-156 + structured_io, io_nohdf5_m.f90:334-577 with
//     TMP_FILE_FORMAT = BINARY): Fortran unformatted sequential files of three
//     records (x, y, z), each n0*n1*n2 reals, with 4-byte record markers.
//
// Pure host code: no HIP calls are made here.
#pragma once
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace lcs {
namespace io {

// Fortran ESw.d edit descriptor as gfortran writes it: one nonzero digit
// before the point, d after, exponent E+zz for |e| <= 99 and +zzz for
// 99 < |e| <= 999 (the letter dropped), right-justified in w characters; a
// minus sign for negative values including -0; Infinity / NaN right-justified;
// w asterisks when the number does not fit.  Writes exactly w characters (no
// terminator).  Digits are the correctly rounded decimal value (round half to
// even), as printf produces them.
void format_es(double v, int w, int d, char* out);
// ES14.6, the output format of write_data_from_master (ASCII_FMT,
// io_nohdf5_m.f90:43): same characters as format_es(v, 14, 6, out), by a
// faster path that falls back to format_es near rounding ties.
void format_es14_6(double v, char* out);

// The file-index part of write_lcs's file name (io_nohdf5_m.f90:71-115): the
// integer in as many digits as it has, except 1000..9999, which the reference
// writes with i10.10 (ten digits, zero-padded).  False when |findex| is out of
// the reference's range (> 999999999).
bool findex_string(long long findex, std::string& out);
// Fortran i4.4 (zero-padded to 4 digits; **** when it does not fit)
std::string i4(long long v);

// One variable (Tecplot block) or one record (unformatted file): n values,
// float or double, kept alive by the job's owners.
struct Array {
  const void* p = nullptr;
  size_t n = 0;
  bool is_double = true;
};

struct Job {
  enum Kind { kTecplot = 0, kUnformatted = 1 } kind = kTecplot;
  std::string path;
  std::string header;                        // Tecplot header lines
  std::vector<Array> arrays;                 // blocks / records in order
  std::vector<std::shared_ptr<void>> keep;   // owners of the arrays' memory
};

// Formats values [first, first + count) of a, 10 per line, each line ended by
// '\n' (first must be a multiple of 10).  Returns the bytes written to out,
// which must hold count*14 + count/10 + 1 bytes.
size_t format_lines(const Array& a, size_t first, size_t count, char* out);

// Synchronous writers, used by the background thread.  The file is written
// under path + ".part" and renamed when complete, so that a reader never sees
// a partial file.  Return false with a message on an I/O failure.
bool write_tecplot(const Job& j, int nthreads, std::string& err);
bool write_unformatted(const Job& j, std::string& err);

// Creates directory dir (one level, as "mkdir -p ./dir" of cfd2lcs.f90:35-36
// does for the fixed names) if it does not exist.
bool ensure_dir(const std::string& dir, std::string& err);

// FIFO of jobs run by one background thread.  submit() blocks while
// max_pending jobs are waiting; flush() waits until every submitted job is
// written.  The first failure is kept and reported by failed().
class Writer {
 public:
  explicit Writer(int max_pending = 2);
  ~Writer();
  void submit(Job&& j);
  void flush();
  bool failed(std::string* msg);

 private:
  void run();
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<Job> q_;
  int busy_ = 0;
  bool stop_ = false;
  bool err_ = false;
  std::string msg_;
  int max_pending_;
  int nthreads_;
  std::thread th_;
};

}  // namespace io
}  // namespace lcs
