// Rutherford-Boeing / Harwell-Boeing input, restating ReadHB_PARA
// (include/sympack/utility.hpp:1274-1482) for one rank: four header lines,
// fixed-width "(nIw)" / "(nEw.d)" records, and byte-counted reads that count
// one byte per line break.  The writer emits exactly that layout.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

#include "analysis.hpp"

namespace spk {

void read_rb(const std::string& filename, CscMatrix& out) {
  std::ifstream infile(filename.c_str(), std::ios::binary);
  if (!infile) throw std::runtime_error("cannot open matrix file " + filename);
  std::string line;
  std::stringstream iss;
  std::getline(infile, line);  // title/key
  Int colptrCnt = 0;
  Ptr rowindCnt = 0, nzvalCnt = 0;
  if (std::getline(infile, line)) {
    iss.str(line);
    iss.clear();
    Ptr dummy;
    iss >> dummy >> colptrCnt >> rowindCnt >> nzvalCnt;
  }
  Int m = 0, n = 0;
  Ptr nnz = 0;
  if (std::getline(infile, line)) {
    iss.str(line);
    iss.clear();
    std::string type;
    iss >> type >> m >> n >> nnz;
  }
  int colptrWidth = 0, rowindWidth = 0, nzvalWidth = 0;
  int colptrPerRow = 0, rowindPerRow = 0, nzvalPerRow = 0, dummy = 0;
  if (std::getline(infile, line)) {
    iss.str(line);
    iss.clear();
    std::string fmt;
    iss >> fmt;
    std::sscanf(fmt.c_str(), "(%dI%d)", &colptrPerRow, &colptrWidth);
    iss >> fmt;
    std::sscanf(fmt.c_str(), "(%dI%d)", &rowindPerRow, &rowindWidth);
    iss >> fmt;
    std::sscanf(fmt.c_str(), "(%dE%d.%d)", &nzvalPerRow, &nzvalWidth, &dummy);
  }
  if (n <= 0 || colptrPerRow <= 0 || rowindPerRow <= 0 || nzvalPerRow <= 0)
    throw std::logic_error("unsupported or malformed Harwell-Boeing header in " + filename);

  out.n = n;
  const Idx firstNode = 1;
  const Idx nlocal = n;
  out.colptr.assign(nlocal + 1, 0);
  // colptr block
  {
    size_t lineLastNode = (size_t)std::ceil(double(firstNode + nlocal) / double(colptrPerRow));
    size_t lineFirstNode = (size_t)std::ceil(double(firstNode) / double(colptrPerRow));
    size_t skip = (size_t)(firstNode - 1) * colptrWidth + (lineFirstNode - 1);
    size_t readBytes = (size_t)(nlocal + 1) * colptrWidth + (lineLastNode - lineFirstNode);
    size_t skipAfter = (size_t)(n + 1 - (firstNode + nlocal)) * colptrWidth +
                       (colptrCnt - lineLastNode + 1);
    infile.seekg(skip, std::ios_base::cur);
    std::string rd(readBytes, '\0');
    infile.read(&rd[0], readBytes);
    iss.str(rd);
    iss.clear();
    Ptr v;
    Idx pos = 0;
    while (iss >> v && pos <= nlocal) out.colptr[pos++] = v;
    infile.seekg(skipAfter, std::ios_base::cur);
  }
  Ptr first_idx = out.colptr.front(), last_idx = out.colptr.back();
  for (Idx i = nlocal + 1; i >= 1; i--) out.colptr[i - 1] = out.colptr[i - 1] - out.colptr[0] + 1;
  Ptr nnzLocal = out.colptr.back() - 1;
  out.rowind.assign(nnzLocal, 0);
  {
    size_t lineLastEdge = (size_t)std::ceil(double(last_idx - 1) / double(rowindPerRow));
    size_t lineFirstEdge = (size_t)std::ceil(double(first_idx) / double(rowindPerRow));
    size_t skip = (size_t)(first_idx - 1) * rowindWidth + (lineFirstEdge - 1);
    size_t readBytes = (size_t)(last_idx - first_idx) * rowindWidth + (lineLastEdge - lineFirstEdge);
    size_t skipAfter = (size_t)(nnz + 1 - last_idx) * rowindWidth + (rowindCnt - lineLastEdge + 1);
    infile.seekg(skip, std::ios_base::cur);
    std::string rd(readBytes, '\0');
    infile.read(&rd[0], readBytes);
    iss.str(rd);
    iss.clear();
    Idx v;
    Ptr pos = 0;
    while (iss >> v && pos < nnzLocal) out.rowind[pos++] = v;
    infile.seekg(skipAfter, std::ios_base::cur);
  }
  out.nzval.assign(nnzLocal, 0.0);
  {
    size_t lineLastEdge = (size_t)std::ceil(double(last_idx - 1) / double(nzvalPerRow));
    size_t lineFirstEdge = (size_t)std::ceil(double(first_idx) / double(nzvalPerRow));
    size_t skip = (size_t)(first_idx - 1) * nzvalWidth + (lineFirstEdge - 1);
    size_t readBytes = (size_t)(last_idx - first_idx) * nzvalWidth + (lineLastEdge - lineFirstEdge);
    infile.seekg(skip, std::ios_base::cur);
    std::string rd(readBytes, '\0');
    infile.read(&rd[0], readBytes);
    iss.str(rd);
    iss.clear();
    double v;
    Ptr pos = 0;
    while (iss >> v && pos < nnzLocal) out.nzval[pos++] = v;
  }
  // sort each column's rows (and values) ascending (utility.hpp:1465-1475)
  for (Idx c = 0; c < nlocal; ++c) {
    Ptr b = out.colptr[c] - 1, e = out.colptr[c + 1] - 1;
    std::vector<std::pair<Idx, double>> tmp;
    tmp.reserve(e - b);
    for (Ptr p = b; p < e; ++p) tmp.emplace_back(out.rowind[p], out.nzval[p]);
    std::stable_sort(tmp.begin(), tmp.end(),
                     [](const std::pair<Idx, double>& a, const std::pair<Idx, double>& b2) {
                       return a.first < b2.first;
                     });
    for (Ptr p = b; p < e; ++p) {
      out.rowind[p] = tmp[p - b].first;
      out.nzval[p] = tmp[p - b].second;
    }
  }
}

void write_rb(const std::string& filename, const CscMatrix& a, const std::string& title,
              const std::string& key) {
  const int ipr = 10, iw = 10;  // (10I10)
  const int vpr = 3, vw = 26;   // (3E26.16): 17 significant digits, exact round trip
  Int n = a.n;
  Ptr nnz = a.colptr[n] - 1;
  auto lines = [](Ptr count, int per) { return (count + per - 1) / per; };
  Ptr ptrcrd = lines(n + 1, ipr), indcrd = lines(nnz, ipr), valcrd = lines(nnz, vpr);
  FILE* f = std::fopen(filename.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + filename);
  std::fprintf(f, "%-72.72s%-8.8s\n", title.c_str(), key.c_str());
  std::fprintf(f, "%14lld%14lld%14lld%14lld%14d\n", (long long)(ptrcrd + indcrd + valcrd),
               (long long)ptrcrd, (long long)indcrd, (long long)valcrd, 0);
  std::fprintf(f, "%-3s%11s%14d%14d%14lld%14d\n", "rsa", "", n, n, (long long)nnz, 0);
  std::fprintf(f, "%-16s%-16s%-20s\n", "(10I10)", "(10I10)", "(3E26.16)");
  for (Ptr i = 0; i <= n; ++i) {
    std::fprintf(f, "%*lld", iw, (long long)a.colptr[i]);
    if ((i + 1) % ipr == 0 || i == n) std::fputc('\n', f);
  }
  for (Ptr i = 0; i < nnz; ++i) {
    std::fprintf(f, "%*d", iw, a.rowind[i]);
    if ((i + 1) % ipr == 0 || i == nnz - 1) std::fputc('\n', f);
  }
  for (Ptr i = 0; i < nnz; ++i) {
    std::fprintf(f, "%*.16E", vw, a.nzval[i]);
    if ((i + 1) % vpr == 0 || i == nnz - 1) std::fputc('\n', f);
  }
  std::fclose(f);
}

}  // namespace spk

namespace spk {

// ParaReadDistSparseMatrix (include/sympack/utility.hpp:1037-1265) on one
// rank: the binary CSC file holds int32 n, int32 nnz, then three records,
// each an int32 count followed by its array: colptr (n+1 int32, 1-based),
// rowind (nnz int32, 1-based), nzval (nnz doubles); the reader skips the
// counts (its byte offsets step over them).  Rows are sorted within each
// column, values following (utility.hpp:1236-1244).
void read_csc_binary(const std::string& filename, CscMatrix& out) {
  std::ifstream f(filename.c_str(), std::ios::binary);
  if (!f) throw std::logic_error("File cannot be opened!");
  auto rd = [&](void* p, size_t bytes, const char* what) {
    f.read((char*)p, (std::streamsize)bytes);
    if (!f) throw std::logic_error(std::string("error reading ") + what);
  };
  int32_t n = 0, nnz = 0, cnt = 0;
  rd(&n, 4, "header");
  rd(&nnz, 4, "header");
  if (n <= 0 || nnz < 0) throw std::logic_error("error reading header");
  rd(&cnt, 4, "colptr");
  std::vector<int32_t> cp(n + 1), ri(nnz);
  rd(cp.data(), 4 * (size_t)(n + 1), "colptr");
  rd(&cnt, 4, "rowind");
  if (cp[0] < 0 || (int64_t)cp[n] - cp[0] != nnz) throw std::logic_error("error reading colptr");
  rd(ri.data(), 4 * (size_t)nnz, "rowind");
  rd(&cnt, 4, "nzval");
  std::vector<double> v(nnz);
  rd(v.data(), 8 * (size_t)nnz, "nzval");
  out.n = n;
  out.colptr.assign(n + 1, 0);
  for (Int c = 0; c <= n; ++c) out.colptr[c] = (Ptr)cp[c] - cp[0] + 1;  // local references
  out.rowind.assign(ri.begin(), ri.end());
  out.nzval.swap(v);
  for (Int c = 0; c < n; ++c) {
    const Ptr b = out.colptr[c] - 1, e = out.colptr[c + 1] - 1;
    if (b > e || e > nnz) throw std::logic_error("error reading colptr");
    std::vector<std::pair<Idx, double>> col;
    col.reserve(e - b);
    for (Ptr q = b; q < e; ++q) {
      if (out.rowind[q] < 1 || out.rowind[q] > n) throw std::logic_error("error reading rowind");
      col.push_back({out.rowind[q], out.nzval[q]});
    }
    std::stable_sort(col.begin(), col.end(),
                     [](const std::pair<Idx, double>& a, const std::pair<Idx, double>& b) { return a.first < b.first; });
    for (Ptr q = b; q < e; ++q) {
      out.rowind[q] = col[q - b].first;
      out.nzval[q] = col[q - b].second;
    }
  }
}

void write_csc_binary(const std::string& filename, const CscMatrix& a) {
  std::ofstream f(filename.c_str(), std::ios::binary);
  if (!f) throw std::runtime_error("cannot write " + filename);
  const int32_t n = a.n, nnz = (int32_t)a.rowind.size(), np1 = n + 1;
  std::vector<int32_t> cp(a.colptr.begin(), a.colptr.end()), ri(a.rowind.begin(), a.rowind.end());
  f.write((const char*)&n, 4);
  f.write((const char*)&nnz, 4);
  f.write((const char*)&np1, 4);
  f.write((const char*)cp.data(), 4 * (size_t)np1);
  f.write((const char*)&nnz, 4);
  f.write((const char*)ri.data(), 4 * (size_t)nnz);
  f.write((const char*)&nnz, 4);
  f.write((const char*)a.nzval.data(), 8 * (size_t)nnz);
  if (!f) throw std::runtime_error("cannot write " + filename);
}

// ReadDistSparseMatrixFormatted (utility.hpp:1486-1688) on one rank: a text
// file "n dummy / nnz dummy / colptr (n+1) / rowind (nnz) / nzval (nnz)",
// whitespace separated, 1-based.  A file holding entries above the diagonal
// is taken as the full symmetric matrix and reduced to its lower triangle
// (DistSparseMatrix::ToLowerTriangular, impl:310-340): entries with
// row >= col are kept in file order.
void read_matrix_formatted(const std::string& filename, CscMatrix& out) {
  std::ifstream f(filename.c_str());
  if (!f.good()) throw std::logic_error("File cannot be openeded!");
  Int n = 0, dummy = 0;
  Ptr nnz = 0;
  f >> n >> dummy >> nnz >> dummy;
  if (!f || n <= 0 || nnz < 0) throw std::logic_error("malformed matrix header in " + filename);
  std::vector<Ptr> cp(n + 1);
  for (Int i = 0; i <= n; ++i) f >> cp[i];
  if (!f || cp[0] < 1 || cp[n] - cp[0] != nnz) throw std::logic_error("malformed colptr in " + filename);
  std::vector<Idx> ri(nnz);
  bool lower = true;
  for (Int c = 1; c <= n; ++c)
    for (Ptr q = cp[c - 1]; q < cp[c]; ++q) {
      f >> ri[q - cp[0]];
      if (ri[q - cp[0]] < c) lower = false;
    }
  std::vector<double> v(nnz);
  for (Ptr q = 0; q < nnz; ++q) f >> v[q];
  if (!f) throw std::logic_error("malformed matrix data in " + filename);
  for (Ptr q = 0; q < nnz; ++q)
    if (ri[q] < 1 || ri[q] > n) throw std::logic_error("row index out of range in " + filename);
  out.n = n;
  out.colptr.assign(n + 1, 1);
  out.rowind.clear();
  out.nzval.clear();
  for (Int c = 1; c <= n; ++c) {
    for (Ptr q = cp[c - 1] - cp[0]; q < cp[c] - cp[0]; ++q)
      if (lower || ri[q] >= c) {
        out.rowind.push_back(ri[q]);
        out.nzval.push_back(v[q]);
      }
    out.colptr[c] = (Ptr)out.rowind.size() + 1;
  }
}

void write_matrix_formatted(const std::string& filename, const CscMatrix& a) {
  std::FILE* f = std::fopen(filename.c_str(), "w");
  if (!f) throw std::runtime_error("cannot write " + filename);
  const Ptr nnz = (Ptr)a.rowind.size();
  std::fprintf(f, "%d %d\n%lld %d\n", a.n, a.n, (long long)nnz, 0);
  for (Int i = 0; i <= a.n; ++i) std::fprintf(f, "%lld\n", (long long)a.colptr[i]);
  for (Ptr q = 0; q < nnz; ++q) std::fprintf(f, "%d\n", a.rowind[q]);
  for (Ptr q = 0; q < nnz; ++q) std::fprintf(f, "%.17g\n", a.nzval[q]);
  if (std::fclose(f) != 0) throw std::runtime_error("cannot write " + filename);
}

// ReadMatrix's dispatch (utility.hpp:1702-1716)
void read_matrix(const std::string& filename, const std::string& format, CscMatrix& out) {
  if (format == "CSC")
    read_csc_binary(filename, out);
  else if (format == "HB" || format == "RB" || format == "HARWELL_BOEING")
    read_rb(filename, out);
  else if (format == "matrix")
    read_matrix_formatted(filename, out);
  else
    throw std::logic_error("Unknown matrix format.");
}

}  // namespace spk
