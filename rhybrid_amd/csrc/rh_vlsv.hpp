// rh_vlsv.hpp -- host-side writer of VLSV files (the format of the
// reference's output, src/operator_userdata.cpp:55-311, as the reference's
// reader tools/rhybridpy.py:113-126 reads it): a 16-byte header whose second
// 8 bytes hold the byte offset of an XML footer, the raw little-endian arrays,
// then the footer <VLSV>...</VLSV> with one element per array whose text is
// the array's byte offset and whose attributes give name, mesh, arraysize,
// vectorsize, datatype (float / int / uint) and datasize (bytes).
#pragma once

#include <cstdint>
#include <cstdio>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>

namespace rh {

class VlsvWriter {
public:
    explicit VlsvWriter(const std::string& path) : path_(path) {
        f_ = std::fopen(path.c_str(), "wb");
        if (!f_) throw std::runtime_error("cannot open " + path);
        const uint64_t zero[2] = {0, 0};
        put(zero, sizeof(zero));
        offset_ = sizeof(zero);
    }
    ~VlsvWriter() {
        if (f_) std::fclose(f_);
    }
    VlsvWriter(const VlsvWriter&) = delete;
    VlsvWriter& operator=(const VlsvWriter&) = delete;

    // One array of arraysize x vectorsize elements of `datasize` bytes.
    void array(const std::string& tag, const std::map<std::string, std::string>& attribs, const void* data,
               uint64_t arraysize, uint64_t vectorsize, const char* datatype, uint64_t datasize) {
        std::ostringstream e;
        e << "<" << tag;
        for (const auto& kv : attribs) e << " " << kv.first << "=\"" << escape(kv.second) << "\"";
        e << " arraysize=\"" << arraysize << "\" vectorsize=\"" << vectorsize << "\" datatype=\"" << datatype
          << "\" datasize=\"" << datasize << "\">" << offset_ << "</" << tag << ">\n";
        footer_ += e.str();
        const uint64_t bytes = arraysize * vectorsize * datasize;
        if (bytes) put(data, bytes);
        offset_ += bytes;
    }

    void close() {
        if (!f_) return;
        const std::string xml = "<VLSV>\n" + footer_ + "</VLSV>\n";
        const uint64_t at = offset_;
        put(xml.data(), xml.size());
        if (std::fseek(f_, 8, SEEK_SET) != 0) throw std::runtime_error("seek failed in " + path_);
        put(&at, sizeof(at));
        if (std::fclose(f_) != 0) { f_ = nullptr; throw std::runtime_error("close failed for " + path_); }
        f_ = nullptr;
    }

private:
    void put(const void* p, size_t n) {
        if (std::fwrite(p, 1, n, f_) != n) throw std::runtime_error("write failed for " + path_);
    }
    static std::string escape(const std::string& s) {
        std::string o;
        for (char c : s) {
            switch (c) {
            case '&': o += "&amp;"; break;
            case '<': o += "&lt;"; break;
            case '>': o += "&gt;"; break;
            case '"': o += "&quot;"; break;
            default: o += c;
            }
        }
        return o;
    }
    std::string path_;
    std::FILE* f_ = nullptr;
    uint64_t offset_ = 0;
    std::string footer_;
};

}  // namespace rh
