// The snapshot writer of the drop-in programs under ThreadSanitizer: frames appended from the caller's
// thread while the worker thread writes the previous ones (the queue holds two), buffers recycled through
// buffer(), then finish().  Exits non-zero if a call fails or the file is missing.
#include <cstdio>
#include <string>
#include <vector>

#include "snapshot_writer.hpp"

int main(int argc, char** argv) {
    const std::string path = std::string(argc > 1 ? argv[1] : "/tmp") + "/tsan_snap.h5";
    ccprog::SnapshotWriter w;
    const hsize_t ny = 24, nx = 40;
    if (!w.open(path, {"/a", "/b", "/c"}, {ny, nx})) {
        std::fprintf(stderr, "open: %s\n", w.error().c_str());
        return 1;
    }
    for (int f = 0; f < 40; ++f) {
        std::vector<double> buf = w.buffer();
        for (size_t q = 0; q < buf.size(); ++q) buf[q] = f + 1e-3 * (double)q;
        if (!w.append(0.1 * f, std::move(buf))) {
            std::fprintf(stderr, "append: %s\n", w.error().c_str());
            return 1;
        }
    }
    std::vector<double> x(nx), y(ny);
    if (!w.finish({{"/x", x}, {"/y", y}}) || w.frames() != 40) {
        std::fprintf(stderr, "finish: %s\n", w.error().c_str());
        return 1;
    }
    if (std::FILE* fh = std::fopen(path.c_str(), "rb")) {
        std::fclose(fh);
    } else {
        return 1;
    }
    std::remove(path.c_str());
    std::printf("snapshot writer under TSan: ok\n");
    return 0;
}
