// VTU snapshots of the uniform grid (reference core/src/flow/case.cpp:175-241).
//
// ASCII: the reference's layout byte for byte.  Its doubles are printed by an
// ostream with precision 8 and the default float field, i.e. printf's "%.8g",
// which is what std::to_chars' general format with precision 8 produces; the
// text is formatted in parallel chunks and written in one pass.  Binary: the
// same arrays in VTK's appended raw encoding (UInt64 byte-count headers, as
// the header_type attribute already declares): full precision, no formatting,
// about the ASCII file's size.  SnapshotWriter takes host copies of a step's
// cell fields and writes them on a thread of its own while the run continues.
#include <algorithm>
#include <atomic>
#include <charconv>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <deque>
#include <exception>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "nabla/flow/case.hpp"

namespace nabla::flow {
namespace {

inline char* putD(char* p, double x) {
  return std::to_chars(p, p + 32, x, std::chars_format::general, 8).ptr;
}
inline char* putU(char* p, unsigned long long x) { return std::to_chars(p, p + 24, x).ptr; }

unsigned textThreads() { return std::clamp(std::thread::hardware_concurrency(), 1u, 16u); }

// Appends the text of items [0, n), item i written by fmt(p, i) (at most
// maxBytes each), formatted in parallel chunks and joined in order.
template <class F>
void parallelText(std::string& out, size_t n, size_t maxBytes, F fmt) {
  const size_t chunks = n < 8192 ? 1 : std::min<size_t>(size_t(textThreads()) * 4, (n + 8191) / 8192);
  std::vector<std::string> parts(chunks);
  std::atomic<size_t> next{0};
  auto work = [&] {
    for (size_t c; (c = next.fetch_add(1)) < chunks;) {
      const size_t a = n * c / chunks, b = n * (c + 1) / chunks;
      std::string& s = parts[c];
      s.resize((b - a) * maxBytes);
      char* p = s.data();
      for (size_t i = a; i < b; ++i) p = fmt(p, i);
      s.resize(size_t(p - s.data()));
    }
  };
  const unsigned T = std::min<size_t>(textThreads(), chunks);
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < T; ++t) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
  size_t total = out.size();
  for (const auto& s : parts) total += s.size();
  out.reserve(total);
  for (const auto& s : parts) out += s;
}

// The eight corner coordinates of cell (i, j, k) in the reference's order.
inline void corners(const UniformGrid& g, size_t c, double (&q)[24]) {
  const size_t nx = size_t(g.nx), ny = size_t(g.ny);
  const int i = int(c % nx), j = int((c / nx) % ny), k = int(c / (nx * ny));
  const double x0 = g.origin.x + i * g.dx, x1 = x0 + g.dx;
  const double y0 = g.origin.y + j * g.dy, y1 = y0 + g.dy;
  const double z0 = g.origin.z + k * g.dz, z1 = z0 + g.dz;
  const double v[24] = {x0, y0, z0, x1, y0, z0, x1, y1, z0, x0, y1, z0,
                        x0, y0, z1, x1, y0, z1, x1, y1, z1, x0, y1, z1};
  std::copy(v, v + 24, q);
}

struct CellArrays {
  std::vector<double> speed, solid;
};
CellArrays derived(const VtuSnapshot& s) {
  const size_t n = s.grid.cellCount();
  CellArrays a;
  a.speed.resize(n);
  a.solid.resize(n);
  for (size_t c = 0; c < n; ++c) {
    a.speed[c] = std::sqrt(s.u[c] * s.u[c] + s.v[c] * s.v[c] + s.w[c] * s.w[c]);
    a.solid[c] = s.grid.solid[c] ? 1.0 : 0.0;
  }
  return a;
}

const char* kHead =
    "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" "
    "byte_order=\"LittleEndian\" header_type=\"UInt64\">\n  <UnstructuredGrid>\n";

std::string asciiText(const VtuSnapshot& s) {
  const UniformGrid& g = s.grid;
  const size_t ncell = g.cellCount(), npts = ncell * 8;
  const CellArrays d = derived(s);
  std::string t = kHead;
  t += "    <Piece NumberOfPoints=\"" + std::to_string(npts) + "\" NumberOfCells=\"" + std::to_string(ncell) + "\">\n";
  t += "      <Points>\n        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n";
  parallelText(t, ncell, 24 * 32, [&](char* p, size_t c) {
    double q[24];
    corners(g, c, q);
    for (int e = 0; e < 24; ++e) {
      p = putD(p, q[e]);
      *p++ = e == 23 ? '\n' : ' ';
    }
    return p;
  });
  t += "        </DataArray>\n      </Points>\n      <Cells>\n";
  t += "        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n";
  parallelText(t, npts, 24, [](char* p, size_t i) {
    p = putU(p, i);
    *p++ = (i + 1) % 8 == 0 ? '\n' : ' ';
    return p;
  });
  t += "        </DataArray>\n        <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n";
  parallelText(t, ncell, 24, [](char* p, size_t i) {
    p = putU(p, (i + 1) * 8);
    *p++ = ' ';
    return p;
  });
  t += "\n        </DataArray>\n        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n";
  parallelText(t, ncell, 3, [](char* p, size_t) {
    *p++ = '1';
    *p++ = '2';
    *p++ = ' ';
    return p;
  });
  t += "\n        </DataArray>\n      </Cells>\n      <CellData Scalars=\"speed\">\n";
  auto dump = [&](const char* name, const std::vector<double>& a) {
    t += "        <DataArray type=\"Float64\" Name=\"";
    t += name;
    t += "\" format=\"ascii\">\n";
    parallelText(t, a.size(), 32, [&](char* p, size_t i) {
      p = putD(p, a[i]);
      *p++ = ' ';
      return p;
    });
    t += "\n        </DataArray>\n";
  };
  dump("u", s.u);
  dump("v", s.v);
  dump("w", s.w);
  dump("p", s.p);
  dump("speed", d.speed);
  dump("solid", d.solid);
  t += "      </CellData>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n";
  return t;
}

void writeAll(const std::string& path, const void* data, size_t bytes, std::FILE* f = nullptr) {
  const bool own = f == nullptr;
  if (own) f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("writeGridVtu: cannot open " + path);
  const size_t w = std::fwrite(data, 1, bytes, f);
  if (own && std::fclose(f) != 0) throw std::runtime_error("writeGridVtu: cannot write " + path);
  if (w != bytes) throw std::runtime_error("writeGridVtu: cannot write " + path);
}

void writeBinary(const VtuSnapshot& s, const std::string& path) {
  const UniformGrid& g = s.grid;
  const size_t ncell = g.cellCount(), npts = ncell * 8;
  const CellArrays d = derived(s);
  const std::vector<double>* cellData[6] = {&s.u, &s.v, &s.w, &s.p, &d.speed, &d.solid};
  const char* names[6] = {"u", "v", "w", "p", "speed", "solid"};
  // block sizes (bytes) in appended order: points, connectivity, offsets, types, six cell arrays
  const size_t sizes[10] = {npts * 3 * 8, npts * 8, ncell * 8, ncell, ncell * 8,
                            ncell * 8,    ncell * 8, ncell * 8, ncell * 8, ncell * 8};
  size_t off[10];
  size_t o = 0;
  for (int b = 0; b < 10; ++b) {
    off[b] = o;
    o += 8 + sizes[b];
  }
  std::string x = kHead;
  x += "    <Piece NumberOfPoints=\"" + std::to_string(npts) + "\" NumberOfCells=\"" + std::to_string(ncell) + "\">\n";
  x += "      <Points>\n        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"appended\" offset=\"" +
       std::to_string(off[0]) + "\"/>\n      </Points>\n      <Cells>\n";
  x += "        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"appended\" offset=\"" + std::to_string(off[1]) +
       "\"/>\n";
  x += "        <DataArray type=\"Int64\" Name=\"offsets\" format=\"appended\" offset=\"" + std::to_string(off[2]) +
       "\"/>\n";
  x += "        <DataArray type=\"UInt8\" Name=\"types\" format=\"appended\" offset=\"" + std::to_string(off[3]) +
       "\"/>\n      </Cells>\n      <CellData Scalars=\"speed\">\n";
  for (int a = 0; a < 6; ++a)
    x += std::string("        <DataArray type=\"Float64\" Name=\"") + names[a] + "\" format=\"appended\" offset=\"" +
         std::to_string(off[4 + a]) + "\"/>\n";
  x += "      </CellData>\n    </Piece>\n  </UnstructuredGrid>\n  <AppendedData encoding=\"raw\">\n   _";
  std::FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("writeGridVtu: cannot open " + path);
  try {
    writeAll(path, x.data(), x.size(), f);
    auto block = [&](size_t bytes, const void* data) {
      const uint64_t h = bytes;
      writeAll(path, &h, 8, f);
      writeAll(path, data, bytes, f);
    };
    std::vector<double> pts(npts * 3);
    for (size_t c = 0; c < ncell; ++c) {
      double q[24];
      corners(g, c, q);
      std::copy(q, q + 24, pts.data() + 24 * c);
    }
    block(sizes[0], pts.data());
    std::vector<int64_t> ints(npts);
    for (size_t i = 0; i < npts; ++i) ints[i] = int64_t(i);
    block(sizes[1], ints.data());
    ints.resize(ncell);
    for (size_t i = 0; i < ncell; ++i) ints[i] = int64_t((i + 1) * 8);
    block(sizes[2], ints.data());
    std::vector<uint8_t> types(ncell, 12);
    block(sizes[3], types.data());
    for (int a = 0; a < 6; ++a) block(sizes[4 + a], cellData[a]->data());
    const char* tail = "\n  </AppendedData>\n</VTKFile>\n";
    writeAll(path, tail, std::char_traits<char>::length(tail), f);
  } catch (...) {
    std::fclose(f);
    throw;
  }
  if (std::fclose(f) != 0) throw std::runtime_error("writeGridVtu: cannot write " + path);
}

}  // namespace

VtuSnapshot captureVtu(const NSSolver& solver) {
  VtuSnapshot s;
  s.grid = solver.grid();
  s.u = solver.cellVelocity(0);
  s.v = solver.cellVelocity(1);
  s.w = solver.cellVelocity(2);
  s.p = solver.cellPressure();
  return s;
}

void writeVtu(const VtuSnapshot& s, const std::string& path, VtuFormat fmt) {
  if (fmt == VtuFormat::Binary) {
    writeBinary(s, path);
    return;
  }
  const std::string t = asciiText(s);
  writeAll(path, t.data(), t.size());
}

void writeGridVtu(const NSSolver& solver, const std::string& path) {
  writeVtu(captureVtu(solver), path, VtuFormat::Ascii);
}

// ---- asynchronous writer ----------------------------------------------------
struct SnapshotWriter::Impl {
  VtuFormat fmt;
  std::mutex mu;
  std::condition_variable cv, idle;
  std::deque<std::pair<VtuSnapshot, std::string>> q;
  bool stop = false, busy = false;
  std::exception_ptr err;
  std::thread th;
  void run() {
    for (;;) {
      std::pair<VtuSnapshot, std::string> job;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return stop || !q.empty(); });
        if (q.empty()) return;
        job = std::move(q.front());
        q.pop_front();
        busy = true;
      }
      try {
        if (!err) writeVtu(job.first, job.second, fmt);
      } catch (...) {
        std::lock_guard<std::mutex> lk(mu);
        if (!err) err = std::current_exception();
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        busy = false;
      }
      idle.notify_all();
    }
  }
};

SnapshotWriter::SnapshotWriter(VtuFormat fmt) : impl_(new Impl) {
  impl_->fmt = fmt;
  impl_->th = std::thread([this] { impl_->run(); });
}

SnapshotWriter::~SnapshotWriter() {
  {
    std::lock_guard<std::mutex> lk(impl_->mu);
    impl_->stop = true;
  }
  impl_->cv.notify_all();
  impl_->th.join();
}

void SnapshotWriter::submit(VtuSnapshot s, std::string path) {
  {
    std::lock_guard<std::mutex> lk(impl_->mu);
    if (impl_->err) std::rethrow_exception(impl_->err);
    impl_->q.emplace_back(std::move(s), std::move(path));
  }
  impl_->cv.notify_all();
}

void SnapshotWriter::drain() {
  std::unique_lock<std::mutex> lk(impl_->mu);
  impl_->idle.wait(lk, [&] { return impl_->q.empty() && !impl_->busy; });
  if (impl_->err) std::rethrow_exception(impl_->err);
}

}  // namespace nabla::flow
