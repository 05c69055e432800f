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
