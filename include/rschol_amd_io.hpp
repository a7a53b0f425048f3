// rschol_amd_io.hpp -- the reference's on-disk formats for C++ callers
// (header-only; link with -lz for the .gz matrix files):
//
//   .bcsm / .bcsm.gz   SPARSE_MATRIX::readFromBinary(GZ)
//                      (src/rschol/linearalgebra/SPARSE_MATRIX.cpp:1557-1718):
//                      uint64 nnz, uint64 nrow, uint64 ncol, int32 p[ncol+1],
//                      int32 i[nnz], float64 x[nnz]
//   .vector            VECTOR::read / write (VECTOR_FAST.cpp:194-276):
//                      int32 n, float64 x[n]
//   meshPositions.dat  readVector<VEC3F> (src/rschol/util/STLUtil.h:105-138):
//                      int32 count, count x 3 float64
//   ordering dumps     written by every reference factorization:
//                      SparseSolver_permutation.vector (SparseSolver.cpp:136,
//                      the final permutation, perm[new] = old),
//                      separatorParents.vector (NestedDissection.cpp:395,
//                      cParent per nested-dissection component) and
//                      separators.vector (NestedDissection.cpp:394: per
//                      component {int size; int first; int last;
//                      int numChildren}, CHOLMOD_Environment.h:233-240), all
//                      in readVector's format (int32 count, then the records)
//
// readReferenceOrdering turns the three dumps into the (perm, cparent,
// cmember) triple that rsc_factor_create_with_ordering takes, so a
// factorization here uses the partition of a reference run.  The dumped
// permutation already holds the reference's separator sub-ordering, so the
// factor is created with reorder_separators = 0 (the diagonal block ranges
// depend only on the separator sizes).
#ifndef RSCHOL_AMD_IO_HPP
#define RSCHOL_AMD_IO_HPP

#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "rschol_amd.hpp"

namespace rschol_amd {

namespace io_detail {

struct Reader {
  gzFile gz = nullptr;  // gzread also reads uncompressed files
  explicit Reader(const std::string& path) { gz = gzopen(path.c_str(), "rb"); }
  ~Reader() {
    if (gz) gzclose(gz);
  }
  bool ok() const { return gz != nullptr; }
  bool read(void* dst, size_t bytes) {
    char* p = (char*)dst;
    while (bytes > 0) {
      const unsigned chunk = bytes > (1u << 30) ? (1u << 30) : (unsigned)bytes;
      const int got = gzread(gz, p, chunk);
      if (got <= 0) return false;
      p += got;
      bytes -= (size_t)got;
    }
    return true;
  }
};

inline void fail(const std::string& what) { throw Error(RSC_BAD_INPUT, what); }

template <typename T>
std::vector<T> read_records(const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) fail("cannot open " + path);
  int32_t count = 0;
  std::vector<T> out;
  bool ok = std::fread(&count, sizeof(count), 1, f) == 1 && count >= 0;
  if (ok) {
    out.resize((size_t)count);
    ok = count == 0 || std::fread(out.data(), sizeof(T), (size_t)count, f) == (size_t)count;
  }
  std::fclose(f);
  if (!ok) fail("truncated file " + path);
  return out;
}

template <typename T>
void write_records(const std::string& path, const std::vector<T>& v) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) fail("cannot create " + path);
  const int32_t count = (int32_t)v.size();
  bool ok = std::fwrite(&count, sizeof(count), 1, f) == 1 &&
            (v.empty() || std::fwrite(v.data(), sizeof(T), v.size(), f) == v.size());
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) fail("cannot write " + path);
}

}  // namespace io_detail

// SPARSE_MATRIX::readFromBinaryGZ (also reads an uncompressed .bcsm).
inline SparseColumnMatrix readFromBinaryGZ(const std::string& path) {
  io_detail::Reader r(path);
  if (!r.ok()) io_detail::fail("cannot open " + path);
  uint64_t hdr[3];
  if (!r.read(hdr, sizeof(hdr))) io_detail::fail("truncated header in " + path);
  const uint64_t nnz = hdr[0], nrow = hdr[1], ncol = hdr[2];
  if (nrow != ncol || nrow > (1u << 31) - 2 || nnz > (1ull << 31) - 1)
    io_detail::fail("unsupported matrix shape in " + path);
  SparseColumnMatrix A;
  A._nrow = (int)nrow;
  A._ncol = (int)ncol;
  A._p.resize(ncol + 1);
  A._i.resize(nnz);
  A._x.resize(nnz);
  if (!r.read(A._p.data(), A._p.size() * sizeof(int)) ||
      !r.read(A._i.data(), A._i.size() * sizeof(int)) ||
      !r.read(A._x.data(), A._x.size() * sizeof(double)))
    io_detail::fail("truncated matrix data in " + path);
  if (A._p[0] != 0 || (uint64_t)A._p[ncol] != nnz)
    io_detail::fail("inconsistent column pointers in " + path);
  return A;
}

// VECTOR::read: int32 n, float64 x[n].
inline VECTOR readVectorFile(const std::string& path) {
  return io_detail::read_records<double>(path);
}
inline void writeVectorFile(const std::string& path, const VECTOR& x) {
  io_detail::write_records(path, x);
}

// readVector<VEC3F>: int32 count, count x 3 float64.  Returns false (and
// leaves pos empty) when the file does not exist, as the reference driver
// treats a missing positions file (src/pcg_solver.cpp:85-96).
inline bool readPositions(const std::string& path, Vector3Array& pos) {
  pos.clear();
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fclose(f);
  pos = io_detail::read_records<VEC3F>(path);
  return true;
}

// One record of separators.vector (CHOLMOD_Environment.h:233-240).
struct SeparatorRecord {
  int32_t size, first, last, numChildren;
};

struct Ordering {
  std::vector<int> perm;     // perm[new] = old
  std::vector<int> cparent;  // per component, -1 for a root
  std::vector<int> cmember;  // per original index: its component
};

// The three reference dumps in directory `dir` -> the CHOLMOD-format
// nested-dissection triple.  cmember is rebuilt from the components'
// column ranges in the permuted order.
inline Ordering readReferenceOrdering(const std::string& dir) {
  const std::string sep = dir.empty() ? std::string() : dir + "/";
  Ordering o;
  o.perm = io_detail::read_records<int>(sep + "SparseSolver_permutation.vector");
  o.cparent = io_detail::read_records<int>(sep + "separatorParents.vector");
  std::vector<SeparatorRecord> seps =
      io_detail::read_records<SeparatorRecord>(sep + "separators.vector");
  const int n = (int)o.perm.size(), nc = (int)o.cparent.size();
  if ((int)seps.size() != nc) io_detail::fail("separator and parent counts differ");
  o.cmember.assign(n, -1);
  int next = 0;
  for (int c = 0; c < nc; ++c) {
    const SeparatorRecord& s = seps[c];
    if (s.first != next || s.last - s.first + 1 != s.size || s.last >= n ||
        o.cparent[c] < -1 || o.cparent[c] >= nc)
      io_detail::fail("separator ranges do not tile the permutation");
    for (int k = s.first; k <= s.last; ++k) {
      const int v = o.perm[k];
      if (v < 0 || v >= n || o.cmember[v] != -1) io_detail::fail("not a permutation");
      o.cmember[v] = c;
    }
    next = s.last + 1;
  }
  if (next != n) io_detail::fail("separator ranges do not cover the permutation");
  return o;
}

// Writes a triple in the reference's dump format (the inverse of
// readReferenceOrdering; components must be contiguous in perm).
inline void writeReferenceOrdering(const std::string& dir, const Ordering& o) {
  const std::string sep = dir.empty() ? std::string() : dir + "/";
  const int n = (int)o.perm.size(), nc = (int)o.cparent.size();
  std::vector<SeparatorRecord> seps(nc, SeparatorRecord{0, 0, -1, 0});
  for (int c = 0; c < nc; ++c)
    if (o.cparent[c] >= 0) seps.at(o.cparent[c]).numChildren += 1;
  int cur = 0;
  for (int k = 0; k < n; ++k) {
    const int c = o.cmember.at(o.perm[k]);
    if (c < cur || c >= nc) io_detail::fail("components are not contiguous in perm");
    if (seps[c].size == 0) seps[c].first = k;
    seps[c].size += 1;
    seps[c].last = k;
    cur = c;
  }
  io_detail::write_records(sep + "SparseSolver_permutation.vector", o.perm);
  io_detail::write_records(sep + "separatorParents.vector", o.cparent);
  io_detail::write_records(sep + "separators.vector", seps);
}

}  // namespace rschol_amd

#endif
