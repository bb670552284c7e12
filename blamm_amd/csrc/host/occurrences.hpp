// The occurrence output of `blamm scan` (pwmscan.cpp:81-102): one line per
// hit, seqName  blamm  motif  pos  pos+L  score  strand  .  . , formatted by
// worker threads and written with pwrite at final offsets.
#pragma once

#include <fcntl.h>
#include <unistd.h>

#include <cerrno>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/blamm_hip.h"
#include "host.hpp"

namespace blamm {

// Name, length and strand of a PWM column, as the occurrence lines print them.
struct ColumnInfo {
    std::string name;
    uint64_t len;
    char strand;
};

// The occurrence file.  Each batch of formatted lines arrives as one buffer
// per worker; the workers' buffers are written at their final offsets by one
// thread each (pwrite), so the copy into the page cache runs in parallel.  A
// destination without offsets (a pipe) is written in order.
class OutFile {
public:
    explicit OutFile(const std::string& path) : path_(path)
    {
        fd_ = ::open(path.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
        if (fd_ < 0) throw std::runtime_error("Cannot write to file: " + path);
        seekable_ = ::lseek(fd_, 0, SEEK_CUR) >= 0;
    }
    ~OutFile()
    {
        if (fd_ >= 0) ::close(fd_);
    }
    OutFile(const OutFile&) = delete;
    OutFile& operator=(const OutFile&) = delete;

    struct Part {
        const char* p = nullptr;
        size_t n = 0;
    };

    void write_parts(const std::vector<Part>& parts)
    {
        if (!seekable_) {
            for (const Part& p : parts) put(p.p, p.n, -1);
            return;
        }
        std::vector<uint64_t> at(parts.size());
        for (size_t i = 0; i < parts.size(); ++i) {
            at[i] = off_;
            off_ += parts[i].n;
        }
        std::vector<std::string> err(parts.size());
        auto one = [&](size_t i) {
            try {
                put(parts[i].p, parts[i].n, (int64_t)at[i]);
            } catch (const std::exception& e) {
                err[i] = e.what();
            }
        };
        std::vector<std::thread> pool;
        for (size_t i = 1; i < parts.size(); ++i) pool.emplace_back(one, i);
        if (!parts.empty()) one(0);
        for (auto& t : pool) t.join();
        for (const std::string& e : err)
            if (!e.empty()) throw std::runtime_error(e);
    }

    // For writers that place parts themselves: whether the file takes
    // offsets, the offset the next part starts at, a write of [p, p + n) at
    // offset `at` (thread-safe) or at the current position (a pipe), and
    // moving the next-part offset past what was written at offsets.
    bool seekable() const { return seekable_; }
    uint64_t offset() const { return off_; }
    void write_at(const char* p, size_t n, uint64_t at) { put(p, n, (int64_t)at); }
    void write_next(const char* p, size_t n) { put(p, n, -1); }
    void advance(uint64_t n) { off_ += n; }

    void close()
    {
        if (fd_ >= 0 && ::close(fd_) != 0) {
            fd_ = -1;
            throw std::runtime_error("Cannot write to file: " + path_);
        }
        fd_ = -1;
    }

private:
    // all of [p, p + n) at offset `at` (or at the current position when at < 0)
    void put(const char* p, size_t n, int64_t at)
    {
        while (n > 0) {
            const ssize_t w = at >= 0 ? ::pwrite(fd_, p, n, (off_t)at) : ::write(fd_, p, n);
            if (w < 0) {
                if (errno == EINTR) continue;
                throw std::runtime_error("Cannot write to file: " + path_ + " (" + std::strerror(errno) + ")");
            }
            p += w;
            n -= (size_t)w;
            if (at >= 0) at += w;
        }
    }

    std::string path_;
    int fd_ = -1;
    bool seekable_ = true;
    uint64_t off_ = 0;
};

// Writes the lines of hits[0..n) whose position is below `payload` (hits
// starting in a segment's halo belong to the next segment); positions are
// relative to stream position `base` of s.  Returns the lines written.
uint64_t write_hits(OutFile& os, const blamm_hit* hits, size_t n, uint64_t payload, const PackedStream& s,
                    uint64_t base, const std::vector<ColumnInfo>& cols, const Group& g, int threads);

}  // namespace blamm
