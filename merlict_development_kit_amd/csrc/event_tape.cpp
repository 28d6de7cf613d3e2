// CORSIKA event tapes on the host (include/mlgpu.h, mlgpu_event_tape_*):
// the tar runs event_tape::Run reads (merlict_corsika/event_tape.cpp:10-45
// with mliEventTapeReader, mli_corsika_EventTape_standalone.c:1265-1468).
//
// A tape is a tar of ddddddddd/RUNH.float32 (273 floats), then per event
// ddddddddd/ddddddddd/EVTH.float32 followed by its Cherenkov bunch blocks
// ddddddddd/ddddddddd/ddddddddd.cer.x8.float32 (n x 8 float32), numbered
// from 1.  As in the reference, a tar header that can not be read (end of
// file, zero block, bad checksum or field) ends the tape, an event header
// that fails its checks ends the run, and the bunches of an event end at
// the first member that is not its next block.
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/mlgpu.h"
#include "scene.hpp"

namespace {

constexpr uint64_t kHeaderBytes = 273 * 4;  // a RUNH or EVTH block
constexpr uint64_t kBunchBytes = 8 * 4;

struct TarHeader {
    std::string name;
    uint64_t size = 0;
};

// an unsigned number of a tar header field: octal digits, a trailing space
// or NUL in either of the last two places allowed (mliTar_field_to_uint)
bool octal_field(const char* field, size_t n, uint64_t& out) {
    char buf[32] = {0};
    std::memcpy(buf, field, n);
    if (buf[n - 1] == ' ') buf[n - 1] = 0;
    if (buf[n - 2] == ' ') buf[n - 2] = 0;
    if (buf[0] == 0 || std::isspace(static_cast<unsigned char>(buf[0]))) return false;
    char* end = nullptr;
    const long long v = std::strtoll(buf, &end, 8);
    if (end != buf + std::strlen(buf) || v < 0) return false;
    out = static_cast<uint64_t>(v);
    return true;
}

// the size field in base 256 (first byte 0x80, the next three 0)
bool base256_size(const unsigned char* f, uint64_t& out) {
    if (f[1] != 0 || f[2] != 0 || f[3] != 0) return false;
    uint64_t v = 0;
    for (int i = 4; i < 12; i++) v = (v << 8) | f[i];
    out = v;
    return true;
}

struct Tar {
    std::ifstream f;
    uint64_t pos = 0, remaining = 0;

    bool read(void* dst, uint64_t n) {
        f.read(static_cast<char*>(dst), static_cast<std::streamsize>(n));
        if (!f) return false;
        pos += n;
        return true;
    }

    // false at the end of the tape or on a header that can not be parsed
    bool next_header(TarHeader& h) {
        unsigned char rh[512];
        if (!read(rh, 512)) return false;
        bool zero = true;
        for (unsigned char c : rh) zero = zero && c == 0;
        if (zero) return false;
        // checksum: 256 (the field as spaces) + every other byte
        uint64_t sum = 256;
        for (int i = 0; i < 148; i++) sum += rh[i];
        for (int i = 156; i < 512; i++) sum += rh[i];
        uint64_t want = 0, mode = 0, owner = 0, mtime = 0;
        if (!octal_field(reinterpret_cast<const char*>(rh + 148), 8, want) || want != sum) return false;
        if (!octal_field(reinterpret_cast<const char*>(rh + 100), 8, mode) ||
            !octal_field(reinterpret_cast<const char*>(rh + 108), 8, owner))
            return false;
        if (rh[124] == 0x80) {
            if (!base256_size(rh + 124, h.size)) return false;
        } else if (!octal_field(reinterpret_cast<const char*>(rh + 124), 12, h.size)) {
            return false;
        }
        if (!octal_field(reinterpret_cast<const char*>(rh + 136), 12, mtime)) return false;
        h.name.assign(reinterpret_cast<const char*>(rh), strnlen(reinterpret_cast<const char*>(rh), 100));
        remaining = h.size;
        return true;
    }

    // n bytes of the current member; its last read also skips the padding
    // to the next 512-byte record
    bool data(void* dst, uint64_t n) {
        if (remaining < n || !read(dst, n)) return false;
        remaining -= n;
        if (remaining == 0) {
            const uint64_t pad = (512 - pos % 512) % 512;
            char skip[512];
            if (pad && !read(skip, pad)) return false;
        }
        return true;
    }
};

// the name against a template whose 'd' stand for digits
bool matches(const std::string& s, const char* t) {
    if (s.size() != std::strlen(t)) return false;
    for (size_t i = 0; i < s.size(); i++) {
        if (t[i] == 'd') {
            if (!std::isdigit(static_cast<unsigned char>(s[i]))) return false;
        } else if (s[i] != t[i]) {
            return false;
        }
    }
    return true;
}

// the nine decimal digits at s[at..]
uint64_t digits9(const std::string& s, size_t at) { return std::strtoull(s.substr(at, 9).c_str(), nullptr, 10); }

float chars_to_float(const char* c) {
    float f;
    std::memcpy(&f, c, 4);
    return f;
}

}  // namespace

struct mlgpu_event_tape {
    Tar tar;
    TarHeader th;
    bool has_th = false;
    uint64_t run_number = 0, event_number = 0, block_number = 0;
    uint64_t block_size = 0, block_at = 0;
    bool bunches_left = false;
    float run_header[273] = {};
    float next_header[273] = {};
    bool has_events = false;
    float header[273] = {};
    std::vector<float> bunches;

    bool block_ok() const {
        return has_th && matches(th.name, "ddddddddd/ddddddddd/ddddddddd.cer.x8.float32") &&
               digits9(th.name, 0) == run_number && digits9(th.name, 10) == event_number &&
               digits9(th.name, 20) == block_number;
    }

    // mliEventTapeReader_read_evth: 1 with the header in next_header, 0 when
    // the next member is not an event header or the event fails its checks
    bool read_evth() {
        if (!has_th || !matches(th.name, "ddddddddd/ddddddddd/EVTH.float32")) return false;
        const uint64_t path_event = digits9(th.name, 10), path_run = digits9(th.name, 0);
        if (th.size != kHeaderBytes || !tar.data(next_header, kHeaderBytes)) return false;
        if (next_header[0] != chars_to_float("EVTH")) return false;
        if (static_cast<uint64_t>(next_header[1]) != path_event || static_cast<uint64_t>(next_header[43]) != path_run)
            return false;
        event_number = path_event;
        block_number = 1;
        has_th = tar.next_header(th);
        if (!block_ok() || th.size % kBunchBytes != 0) return false;
        block_size = th.size / kBunchBytes;
        block_at = 0;
        bunches_left = true;
        return true;
    }

    // mliEventTapeReader_read_cherenkov_bunch
    bool read_bunch(float* b) {
        if (!bunches_left) return false;
        if (block_at == block_size) {
            block_number += 1;
            has_th = tar.next_header(th);
            if (!has_th || !matches(th.name, "ddddddddd/ddddddddd/ddddddddd.cer.x8.float32")) {
                bunches_left = false;
                return false;
            }
            if (!block_ok() || th.size % kBunchBytes != 0) return false;
            block_size = th.size / kBunchBytes;
            block_at = 0;
        }
        if (!tar.data(b, kBunchBytes)) return false;
        block_at += 1;
        return true;
    }
};

namespace {

int event_tape_open(const char* path, mlgpu_event_tape** out) {
    if (!path || !out) {
        mlgpu::set_last_error("null argument");
        return MLGPU_ERR_ARGUMENT;
    }
    *out = nullptr;
    auto* t = new mlgpu_event_tape();
    t->tar.f.open(path, std::ios::binary);
    if (!t->tar.f.is_open()) {
        delete t;
        mlgpu::set_last_error(std::string("event tape: can not open file: ") + path);
        return MLGPU_ERR_ARGUMENT;
    }
    t->has_th = t->tar.next_header(t->th);
    const char* err = nullptr;
    if (!t->has_th) err = "expected a tar header";
    else if (!matches(t->th.name, "ddddddddd/RUNH.float32")) err = "expected the first file to be ddddddddd/RUNH.float32";
    else if (t->th.size != kHeaderBytes) err = "expected RUNH to hold 273 floats";
    else if (!t->tar.data(t->run_header, kHeaderBytes)) err = "can not read RUNH";
    else if (t->run_header[0] != chars_to_float("RUNH")) err = "expected RUNH[0] == 'RUNH'";
    if (!err) {
        t->run_number = digits9(t->th.name, 0);
        // MLI_ROUND: round half down
        const double r = t->run_header[1];
        const double rounded = (r - std::floor(r) > 0.5) ? std::ceil(r) : std::floor(r);
        if (t->run_number != static_cast<uint64_t>(rounded)) err = "run number of RUNH's path and RUNH differ";
    }
    if (err) {
        delete t;
        mlgpu::set_last_error(std::string("event tape: ") + err);
        return MLGPU_ERR_FORMAT;
    }
    t->has_th = t->tar.next_header(t->th);
    t->has_events = t->read_evth();
    *out = t;
    return MLGPU_OK;
}

int event_tape_next_event(mlgpu_event_tape* t, float* header, int64_t* num_bunches) {
    if (!t) {
        mlgpu::set_last_error("null tape");
        return MLGPU_ERR_ARGUMENT;
    }
    if (!t->has_events) {
        mlgpu::set_last_error("event tape: no event left");
        return MLGPU_ERR_ARGUMENT;
    }
    std::memcpy(t->header, t->next_header, sizeof(t->header));
    t->bunches.clear();
    float b[8];
    while (t->read_bunch(b)) t->bunches.insert(t->bunches.end(), b, b + 8);
    t->has_events = t->read_evth();
    if (header) std::memcpy(header, t->header, sizeof(t->header));
    if (num_bunches) *num_bunches = static_cast<int64_t>(t->bunches.size() / 8);
    return MLGPU_OK;
}

}  // namespace

extern "C" {

// A damaged file must not end the caller: allocation failures and other
// exceptions become format errors at the C boundary.
int mlgpu_event_tape_open(const char* path, mlgpu_event_tape** out) {
    try {
        return event_tape_open(path, out);
    } catch (const std::exception& e) {
        mlgpu::set_last_error(std::string("event tape: ") + e.what());
        return MLGPU_ERR_FORMAT;
    }
}

void mlgpu_event_tape_close(mlgpu_event_tape* t) { delete t; }

int mlgpu_event_tape_run_header(const mlgpu_event_tape* t, float* raw) {
    if (!t || !raw) {
        mlgpu::set_last_error("null argument");
        return MLGPU_ERR_ARGUMENT;
    }
    std::memcpy(raw, t->run_header, sizeof(t->run_header));
    return MLGPU_OK;
}

int mlgpu_event_tape_has_events(const mlgpu_event_tape* t) { return t && t->has_events ? 1 : 0; }

int mlgpu_event_tape_next_event(mlgpu_event_tape* t, float* header, int64_t* num_bunches) {
    try {
        return event_tape_next_event(t, header, num_bunches);
    } catch (const std::exception& e) {
        mlgpu::set_last_error(std::string("event tape: ") + e.what());
        return MLGPU_ERR_FORMAT;
    }
}

int mlgpu_event_tape_event_bunches(const mlgpu_event_tape* t, float* bunches) {
    if (!t || !bunches) {
        mlgpu::set_last_error("null argument");
        return MLGPU_ERR_ARGUMENT;
    }
    std::memcpy(bunches, t->bunches.data(), t->bunches.size() * 4);
    return MLGPU_OK;
}

}  // extern "C"
