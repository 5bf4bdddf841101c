// snapshot_writer.cpp -- see snapshot_writer.hpp.
#include "snapshot_writer.hpp"

#include <cstdio>

namespace ccprog {

SnapshotWriter::~SnapshotWriter() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    if (thread_.joinable()) thread_.join();
    for (hid_t d : dsets_)
        if (d >= 0) H5Dclose(d);
    if (tmp_ >= 0) H5Fclose(tmp_);
}

bool SnapshotWriter::fail(const std::string& m) {
    if (err_.empty()) err_ = m;
    return false;
}

bool SnapshotWriter::open(const std::string& path, const std::vector<std::string>& fields,
                          const std::vector<hsize_t>& frame_dims) {
    path_ = path;
    tmp_path_ = path + ".partial";
    fields_ = fields;
    dims_ = frame_dims;
    cells_ = 1;
    for (hsize_t d : dims_) cells_ *= (size_t)d;
    tmp_ = H5Fcreate(tmp_path_.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (tmp_ < 0) return fail("cannot create " + tmp_path_);
    const int rank = (int)dims_.size() + 1;
    std::vector<hsize_t> cur(rank), mx(rank), chunk(rank);
    cur[0] = 0;
    mx[0] = H5S_UNLIMITED;
    chunk[0] = 1;
    for (int r = 1; r < rank; r++) cur[r] = mx[r] = chunk[r] = dims_[r - 1];
    for (const std::string& name : fields_) {
        hid_t sp = H5Screate_simple(rank, cur.data(), mx.data());
        hid_t pl = H5Pcreate(H5P_DATASET_CREATE);
        H5Pset_chunk(pl, rank, chunk.data());
        hid_t ds = H5Dcreate2(tmp_, name.c_str(), H5T_IEEE_F64LE, sp, H5P_DEFAULT, pl, H5P_DEFAULT);
        H5Pclose(pl);
        H5Sclose(sp);
        if (ds < 0) return fail("cannot create dataset " + name + " in " + tmp_path_);
        dsets_.push_back(ds);
    }
    thread_ = std::thread(&SnapshotWriter::worker, this);
    return true;
}

std::vector<double> SnapshotWriter::buffer() {
    std::lock_guard<std::mutex> lk(mu_);
    if (!pool_.empty()) {
        std::vector<double> b = std::move(pool_.back());
        pool_.pop_back();
        return b;
    }
    return std::vector<double>(cells_ * fields_.size());
}

bool SnapshotWriter::append(double t, std::vector<double>&& frame) {
    if (frame.size() != cells_ * fields_.size()) return fail("frame of the wrong size");
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return queue_.size() < 2 || bad_; });
    if (bad_) return false;
    queue_.emplace_back(t, std::move(frame));
    appended_++;
    lk.unlock();
    cv_.notify_all();
    return true;
}

bool SnapshotWriter::write_frame(double t, const std::vector<double>& f) {
    const int rank = (int)dims_.size() + 1;
    const hsize_t k = (hsize_t)written_;
    std::vector<hsize_t> ext(rank), start(rank, 0), count(rank);
    ext[0] = k + 1;
    start[0] = k;
    count[0] = 1;
    for (int r = 1; r < rank; r++) ext[r] = count[r] = dims_[r - 1];
    hid_t mem = H5Screate_simple(rank, count.data(), nullptr);
    bool ok = true;
    for (size_t i = 0; i < dsets_.size() && ok; i++) {
        if (H5Dset_extent(dsets_[i], ext.data()) < 0) { ok = false; break; }
        hid_t fs = H5Dget_space(dsets_[i]);
        H5Sselect_hyperslab(fs, H5S_SELECT_SET, start.data(), nullptr, count.data(), nullptr);
        if (H5Dwrite(dsets_[i], H5T_NATIVE_DOUBLE, mem, fs, H5P_DEFAULT, f.data() + i * cells_) < 0) ok = false;
        H5Sclose(fs);
    }
    H5Sclose(mem);
    if (!ok) return fail("write to " + tmp_path_ + " failed");
    times_.push_back(t);
    written_++;
    return true;
}

void SnapshotWriter::worker() {
    for (;;) {
        std::pair<double, std::vector<double>> item;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return !queue_.empty() || stop_; });
            if (queue_.empty()) return;  // stop_ with nothing left
            item = std::move(queue_.front());
            queue_.pop_front();
        }
        const bool ok = write_frame(item.first, item.second);
        {
            std::lock_guard<std::mutex> lk(mu_);
            pool_.push_back(std::move(item.second));
            if (!ok) bad_ = true;
        }
        cv_.notify_all();
        if (!ok) return;
    }
}

bool SnapshotWriter::finish(const std::vector<Coord>& coords) {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    if (thread_.joinable()) thread_.join();
    if (bad_) return false;

    hid_t out = H5Fcreate(path_.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (out < 0) return fail("cannot create " + path_);
    auto write1 = [&](const std::string& name, const std::vector<double>& a) {
        hsize_t n = (hsize_t)a.size();
        hid_t sp = H5Screate_simple(1, &n, nullptr);
        hid_t ds = H5Dcreate2(out, name.c_str(), H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        herr_t st = (ds < 0) ? -1 : H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.data());
        if (ds >= 0) H5Dclose(ds);
        H5Sclose(sp);
        return st >= 0;
    };
    bool ok = true;
    for (const Coord& c : coords) ok = ok && write1(c.first, c.second);
    ok = ok && write1("/t", times_);

    const int rank = (int)dims_.size() + 1;
    std::vector<hsize_t> full(rank), start(rank, 0), count(rank);
    full[0] = (hsize_t)times_.size();
    count[0] = 1;
    for (int r = 1; r < rank; r++) full[r] = count[r] = dims_[r - 1];
    std::vector<double> buf(cells_);
    hid_t mem = H5Screate_simple(rank, count.data(), nullptr);
    for (size_t i = 0; i < fields_.size() && ok; i++) {
        hid_t sp = H5Screate_simple(rank, full.data(), nullptr);
        hid_t ds = H5Dcreate2(out, fields_[i].c_str(), H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        if (ds < 0) ok = false;
        for (size_t k = 0; k < times_.size() && ok; k++) {
            start[0] = (hsize_t)k;
            hid_t src = H5Dget_space(dsets_[i]);
            H5Sselect_hyperslab(src, H5S_SELECT_SET, start.data(), nullptr, count.data(), nullptr);
            ok = H5Dread(dsets_[i], H5T_NATIVE_DOUBLE, mem, src, H5P_DEFAULT, buf.data()) >= 0;
            H5Sclose(src);
            hid_t dst = H5Dget_space(ds);
            H5Sselect_hyperslab(dst, H5S_SELECT_SET, start.data(), nullptr, count.data(), nullptr);
            ok = ok && H5Dwrite(ds, H5T_NATIVE_DOUBLE, mem, dst, H5P_DEFAULT, buf.data()) >= 0;
            H5Sclose(dst);
        }
        if (ds >= 0) H5Dclose(ds);
        H5Sclose(sp);
    }
    H5Sclose(mem);
    if (H5Fclose(out) < 0) ok = false;
    for (hid_t& d : dsets_) {
        H5Dclose(d);
        d = -1;
    }
    H5Fclose(tmp_);
    tmp_ = -1;
    if (!ok) return fail("writing " + path_ + " failed");
    std::remove(tmp_path_.c_str());
    return true;
}

}  // namespace ccprog
