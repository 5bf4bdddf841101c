// snapshot_writer.hpp -- streaming HDF5 snapshot output of the drop-in programs.
//
// The reference programs keep every snapshot in host memory and write one HDF5 file at exit with
// contiguous H5T_IEEE_F64LE datasets: 1-D coordinates and /t, and one (Nt, ...frame dims) dataset
// per field (M2:881-925, M3:1080-1128, EH:557-572 + 714-731, NS:549-608).  The writer produces the
// same file without holding the history: each frame is appended by a background thread to
// extendible, one-frame-chunked datasets of a temporary file `<path>.partial`; finish() then
// copies them frame by frame into `<path>` in the reference layout and removes the temporary file.
// Host memory stays at a few frames whatever the run length.
#pragma once
#include <hdf5.h>

#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

namespace ccprog {

class SnapshotWriter {
   public:
    using Coord = std::pair<std::string, std::vector<double>>;

    ~SnapshotWriter();

    // fields: dataset names in output order (e.g. "/rho"); frame_dims: dims of one frame
    // (e.g. {Ny, Nx}).  A frame passed to append() is the fields' planes back to back.
    bool open(const std::string& path, const std::vector<std::string>& fields, const std::vector<hsize_t>& frame_dims);

    // A host buffer for the next frame (fields * cells doubles), recycled from written frames.
    std::vector<double> buffer();
    // Queue a frame at time t; returns false after a write error.  Blocks while two frames wait.
    bool append(double t, std::vector<double>&& frame);
    // Write `path`: coords (e.g. {"/x", x}, {"/y", y}) first, then /t, then the fields.
    bool finish(const std::vector<Coord>& coords);

    // Frames appended so far (caller's thread).
    size_t frames() const { return appended_; }
    const std::string& error() const { return err_; }

   private:
    void worker();
    bool write_frame(double t, const std::vector<double>& f);
    bool fail(const std::string& m);

    std::string path_, tmp_path_, err_;
    std::vector<std::string> fields_;
    std::vector<hsize_t> dims_;
    size_t cells_ = 0;
    hid_t tmp_ = -1;
    std::vector<hid_t> dsets_;
    std::vector<double> times_;  // written frames' times (worker thread until finish)
    size_t written_ = 0, appended_ = 0;

    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::pair<double, std::vector<double>>> queue_;
    std::vector<std::vector<double>> pool_;
    bool stop_ = false, bad_ = false;
    std::thread thread_;
};

}  // namespace ccprog
