// CORSIKA EventIO run reader (include/mlgpu.h, mlgpu_eventio_*): the host
// side that feeds mlgpu_corsika_photons, as eventio::Run does for the
// reference's apps (merlict_corsika/eventio.cpp:282-395).
//
// The objects and their order follow the reference: a run header (type
// 1200, a 273-float block), the input card (1212), the telescope positions
// (1201), then per event its header (1202 + telescope offsets 1203), the
// telescope data (1204) holding photon bunches (sub-object 1205), the event
// end (1209), and the run end (1210).  Reading the events is the
// reference's search: at every step try, in this order, a bunch block, a
// 1204 header, an event header, an event end and the run end, stepping back
// over a header of the wrong type (eventio.cpp:366-412).  Quirks kept for
// parity are marked where they occur.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/mlgpu.h"
#include "scene.hpp"

struct mlgpu_eventio_run {
    std::ifstream f;
    bool run_end_found = false;
    float run_header[273] = {};
    std::string input_card;
    std::vector<float> tel_pos;  // ntel x (x, y, z, r)
    float run_end[273] = {};
    // the event handed out by the last mlgpu_eventio_next_event
    float header[273] = {};
    std::vector<float> offsets;  // n x (toff, xoff, yoff, weight)
    std::vector<float> bunches;  // m x 8
    // read ahead: the header last read and the bunches of the coming event
    float next_header[273] = {};
    std::vector<float> next_offsets;
    std::vector<float> next_bunches;
};

namespace {

using mlgpu::set_last_error;

constexpr int32_t kSync = -736130505;  // eventio.cpp:64-67

enum class Rd { ok, wrong_type, no_sync, bad };

struct Head {
    int32_t type = 0, version = 0, id = 0;
    bool user = false, extended = false, only_sub_objects = false;
    int64_t length = 0;
};

bool read_bytes(std::istream& f, void* dst, std::streamsize n) {
    f.read(static_cast<char*>(dst), n);
    return static_cast<bool>(f);
}

void seek(std::istream& f, std::streampos p) {
    f.clear();
    f.seekg(p);
}

// eventio.cpp:9-80.  A top-level header starts with the sync word; a
// sub-object's header has none.
Rd read_head(std::istream& f, bool top_level, Head& h, std::string& err) {
    const std::streampos start = f.tellg();
    int32_t w[4] = {0, 0, 0, 0};
    if (top_level) {
        if (!read_bytes(f, w, 16)) {
            seek(f, start);
            err = "end of file inside an EventIO header";
            return Rd::bad;
        }
        if (w[0] != kSync) {
            seek(f, start);
            return Rd::no_sync;
        }
    } else if (!read_bytes(f, w + 1, 12)) {
        seek(f, start);
        err = "end of file inside an EventIO sub-object header";
        return Rd::bad;
    }
    const uint32_t t = static_cast<uint32_t>(w[1]);
    h.type = static_cast<int32_t>(t & 0xffffu);
    h.version = static_cast<int32_t>((t & 0xfff00000u) >> 20);
    h.user = (t & (1u << 16)) != 0;
    h.extended = (t & (1u << 17)) != 0;
    h.id = w[2];
    h.only_sub_objects = (static_cast<uint32_t>(w[3]) & (1u << 30)) != 0;
    int64_t len = static_cast<int64_t>(static_cast<uint32_t>(w[3]) & 0x3fffffffu);
    if (h.extended) {
        int32_t ext = 0;
        if (!read_bytes(f, &ext, 4)) {
            seek(f, start);
            err = "end of file inside an extended EventIO header";
            return Rd::bad;
        }
        // the reference combines the extension with `&=`, not `|=`
        // (eventio.cpp:53-62); kept for parity
        len &= (static_cast<int64_t>(ext) & 0xfff) << 12;
    }
    h.length = len;
    return Rd::ok;
}

// a top-level header of type `expect`; on another type the stream steps
// back over it (eventio.cpp:336-352)
Rd get_head(std::istream& f, int32_t expect, Head& h, std::string& err) {
    const Rd r = read_head(f, true, h, err);
    if (r != Rd::ok) return r;
    if (h.type != expect) {
        f.seekg(-4 * (h.extended ? 5 : 4), std::ios::cur);
        return Rd::wrong_type;
    }
    return Rd::ok;
}

// int32 273, then 273 floats (eventio.cpp:96-113)
bool read_273(std::istream& f, float* block, std::string& err) {
    int32_t n = 0;
    if (!read_bytes(f, &n, 4) || n != 273) {
        err = "expected a block of 273 floats";
        return false;
    }
    if (!read_bytes(f, block, 273 * 4)) {
        err = "end of file inside a block of 273 floats";
        return false;
    }
    return true;
}

// the run end: int32 3, then the reference reads 273 floats whatever the
// block holds (eventio.cpp:115-131); what the file does not hold stays 0
bool read_run_end(std::istream& f, float* block, std::string& err) {
    int32_t n = 0;
    if (!read_bytes(f, &n, 4) || n != 3) {
        err = "expected the run end block (3)";
        return false;
    }
    std::memset(block, 0, 273 * 4);
    f.read(reinterpret_cast<char*>(block), 273 * 4);
    f.clear();
    return true;
}

// eventio.cpp:146-171: ntel, then ntel records of (x, y, z, r) read as
// consecutive records (the reference's reading; the same as EventIO's four
// arrays for one telescope)
bool read_tel_pos(std::istream& f, const Head& h, std::vector<float>& out, std::string& err) {
    int32_t ntel = 0;
    if (!read_bytes(f, &ntel, 4) || ntel <= 0 || (h.length - 4) / ntel / 4 != 4) {
        err = "bad telescope positions";
        return false;
    }
    out.resize(static_cast<size_t>(ntel) * 4);
    if (!read_bytes(f, out.data(), static_cast<std::streamsize>(out.size() * 4))) {
        err = "end of file inside the telescope positions";
        return false;
    }
    return true;
}

// eventio.cpp:184-239: narray, toff, xoff[narray], yoff[narray] and, when
// the block holds a third array, the weights (else 1)
bool read_offsets(std::istream& f, const Head& h, std::vector<float>& out, std::string& err) {
    int32_t narray = 0;
    float toff = 0.0f;
    if (!read_bytes(f, &narray, 4) || narray <= 0 || !read_bytes(f, &toff, 4)) {
        err = "bad telescope offsets";
        return false;
    }
    // the object must hold at least the two arrays its count announces
    if (static_cast<int64_t>(narray) * 8 > h.length - 8) {
        err = "telescope offsets longer than their object";
        return false;
    }
    const size_t na = static_cast<size_t>(narray);
    std::vector<float> x(na), y(na), w(na, 1.0f);
    if (!read_bytes(f, x.data(), static_cast<std::streamsize>(na * 4)) ||
        !read_bytes(f, y.data(), static_cast<std::streamsize>(na * 4))) {
        err = "end of file inside the telescope offsets";
        return false;
    }
    const int64_t arrays = (h.length - 8) / narray / 4;
    if (arrays == 3) {
        if (!read_bytes(f, w.data(), static_cast<std::streamsize>(na * 4))) {
            err = "end of file inside the telescope offsets";
            return false;
        }
    } else if (arrays != 2) {
        err = "telescope offsets hold " + std::to_string(arrays) + " arrays, only 2 or 3 are allowed";
        return false;
    }
    out.resize(na * 4);
    for (size_t i = 0; i < na; i++) {
        out[4 * i + 0] = toff;
        out[4 * i + 1] = x[i];
        out[4 * i + 2] = y[i];
        out[4 * i + 3] = w[i];
    }
    return true;
}

// a bunch block (sub-object 1205, eventio.cpp:241-286): array, telescope,
// photons, number of bunches; then the bunches as float32 or, in the
// compact form (version / 1000 == 1), as int16 scaled like the reference
Rd read_bunches(std::istream& f, std::vector<float>& out, std::string& err) {
    Head h;
    const Rd r = read_head(f, false, h, err);
    if (r != Rd::ok) return r;
    if (h.type != 1205) {
        f.seekg(-4 * (h.extended ? 4 : 3), std::ios::cur);
        return Rd::wrong_type;
    }
    int16_t array_tel[2];
    float photons = 0.0f;
    int32_t nb = 0;
    if (!read_bytes(f, array_tel, 4) || !read_bytes(f, &photons, 4) || !read_bytes(f, &nb, 4) || nb < 0) {
        err = "bad photon bunch header";
        return Rd::bad;
    }
    const bool compact = h.version / 1000 == 1;
    // nb bunches of 8 values (2 or 4 bytes each) after the 12-byte head
    if (static_cast<int64_t>(nb) * 8 * (compact ? 2 : 4) > h.length - 12) {
        err = "photon bunches longer than their object";
        return Rd::bad;
    }
    const size_t n = static_cast<size_t>(nb) * 8;
    out.resize(n);
    if (compact) {
        std::vector<int16_t> raw(n);
        if (!read_bytes(f, raw.data(), static_cast<std::streamsize>(n * 2))) {
            err = "end of file inside photon bunches";
            return Rd::bad;
        }
        for (size_t b = 0; b < static_cast<size_t>(nb); b++) {
            float* v = out.data() + 8 * b;
            for (int j = 0; j < 8; j++) v[j] = static_cast<float>(raw[8 * b + j]);
            // the reference's float <- double arithmetic, term by term
            v[0] = static_cast<float>(v[0] * 0.1);
            v[1] = static_cast<float>(v[1] * 0.1);
            v[2] = static_cast<float>(v[2] / 30000.);
            v[3] = static_cast<float>(v[3] / 30000.);
            v[4] = static_cast<float>(v[4] * 0.1);
            v[5] = static_cast<float>(std::pow(10, v[5] * 0.001));
            v[6] = static_cast<float>(v[6] * 0.01);
        }
    } else if (!read_bytes(f, out.data(), static_cast<std::streamsize>(n * 4))) {
        err = "end of file inside photon bunches";
        return Rd::bad;
    }
    return Rd::ok;
}

// the search for the coming event's bunches (eventio.cpp:366-412): fills
// next_header / next_offsets on the way; at the run end the bunches are
// empty and run_end_found is set
bool advance(mlgpu_eventio_run* run, std::string& err) {
    run->next_bunches.clear();
    std::istream& f = run->f;
    while (!run->run_end_found) {
        bool found = false;
        Rd r = read_bunches(f, run->next_bunches, err);
        if (r == Rd::ok) return true;
        if (r == Rd::bad) return false;
        Head h;
        r = get_head(f, 1204, h, err);
        if (r == Rd::ok) found = true;
        else if (r == Rd::bad) return false;
        r = get_head(f, 1202, h, err);
        if (r == Rd::ok) {
            if (!read_273(f, run->next_header, err)) return false;
            // a 1202 not followed by 1203 leaves the stream after the block,
            // as in the reference
            r = get_head(f, 1203, h, err);
            if (r == Rd::ok) {
                if (!read_offsets(f, h, run->next_offsets, err)) return false;
                found = true;
            } else if (r == Rd::bad) {
                return false;
            }
        } else if (r == Rd::bad) {
            return false;
        }
        r = get_head(f, 1209, h, err);
        if (r == Rd::ok) {
            float end[273];
            if (!read_273(f, end, err)) return false;
            found = true;
        } else if (r == Rd::bad) {
            return false;
        }
        r = get_head(f, 1210, h, err);
        if (r == Rd::ok) {
            if (!read_run_end(f, run->run_end, err)) return false;
            found = true;
            run->run_end_found = true;
        } else if (r == Rd::bad) {
            return false;
        }
        if (!found) {
            err = "not a single valid EventIO structure found";
            return false;
        }
    }
    return true;
}

int fail(const std::string& what) {
    set_last_error("eventio: " + what);
    return MLGPU_ERR_FORMAT;
}

}  // namespace

namespace {

int eventio_open(const char* path, mlgpu_eventio_run** out) {
    if (!path || !out) {
        set_last_error("null argument");
        return MLGPU_ERR_ARGUMENT;
    }
    *out = nullptr;
    auto* run = new mlgpu_eventio_run();
    run->f.open(path, std::ios::binary);
    if (!run->f.is_open()) {
        delete run;
        set_last_error(std::string("eventio: can not open file: ") + path);
        return MLGPU_ERR_ARGUMENT;
    }
    std::string err;
    Head h;
    Rd r = get_head(run->f, 1200, h, err);
    bool ok = r == Rd::ok && read_273(run->f, run->run_header, err);
    if (ok) {
        r = get_head(run->f, 1212, h, err);
        ok = r == Rd::ok;
        if (ok && (h.length < 0 || h.length > (int64_t{1} << 26))) {
            err = "input card longer than 64 MiB";
            ok = false;
        }
        if (ok) {
            run->input_card.resize(static_cast<size_t>(h.length));
            ok = h.length == 0 || read_bytes(run->f, &run->input_card[0], static_cast<std::streamsize>(h.length));
            if (!ok) err = "end of file inside the input card";
        }
    }
    if (ok) {
        r = get_head(run->f, 1201, h, err);
        ok = r == Rd::ok && read_tel_pos(run->f, h, run->tel_pos, err);
    }
    if (ok) ok = advance(run, err);
    if (!ok) {
        if (err.empty()) err = r == Rd::no_sync ? "header 'sync' field not correct" : "unexpected object in the run header";
        delete run;
        return fail(err);
    }
    *out = run;
    return MLGPU_OK;
}

int eventio_next_event(mlgpu_eventio_run* run, float* header, int32_t* num_offsets, int64_t* num_bunches) {
    if (!run) {
        set_last_error("null run");
        return MLGPU_ERR_ARGUMENT;
    }
    if (run->run_end_found) {
        set_last_error("eventio: no event left in the run");
        return MLGPU_ERR_ARGUMENT;
    }
    std::memcpy(run->header, run->next_header, sizeof(run->header));
    run->offsets = run->next_offsets;
    run->bunches.swap(run->next_bunches);
    std::string err;
    if (!advance(run, err)) return fail(err);
    if (header) std::memcpy(header, run->header, sizeof(run->header));
    if (num_offsets) *num_offsets = static_cast<int32_t>(run->offsets.size() / 4);
    if (num_bunches) *num_bunches = static_cast<int64_t>(run->bunches.size() / 8);
    return MLGPU_OK;
}

}  // namespace

extern "C" {

// A damaged file must not end the caller: allocation failures and other
// exceptions become format errors at the C boundary.
int mlgpu_eventio_open(const char* path, mlgpu_eventio_run** out) {
    try {
        return eventio_open(path, out);
    } catch (const std::exception& e) {
        return fail(std::string("eventio: ") + e.what());
    }
}

void mlgpu_eventio_close(mlgpu_eventio_run* run) { delete run; }

int mlgpu_eventio_run_header(const mlgpu_eventio_run* run, float* raw, int32_t* num_telescopes,
                             int64_t* input_card_length) {
    if (!run) {
        set_last_error("null run");
        return MLGPU_ERR_ARGUMENT;
    }
    if (raw) std::memcpy(raw, run->run_header, sizeof(run->run_header));
    if (num_telescopes) *num_telescopes = static_cast<int32_t>(run->tel_pos.size() / 4);
    if (input_card_length) *input_card_length = static_cast<int64_t>(run->input_card.size());
    return MLGPU_OK;
}

int mlgpu_eventio_run_info(const mlgpu_eventio_run* run, float* telescope_positions, char* input_card) {
    if (!run) {
        set_last_error("null run");
        return MLGPU_ERR_ARGUMENT;
    }
    if (telescope_positions) std::memcpy(telescope_positions, run->tel_pos.data(), run->tel_pos.size() * 4);
    if (input_card) std::memcpy(input_card, run->input_card.data(), run->input_card.size());
    return MLGPU_OK;
}

int mlgpu_eventio_has_events(const mlgpu_eventio_run* run) { return run && !run->run_end_found ? 1 : 0; }

int mlgpu_eventio_next_event(mlgpu_eventio_run* run, float* header, int32_t* num_offsets, int64_t* num_bunches) {
    try {
        return eventio_next_event(run, header, num_offsets, num_bunches);
    } catch (const std::exception& e) {
        return fail(std::string("eventio: ") + e.what());
    }
}

int mlgpu_eventio_event_data(const mlgpu_eventio_run* run, float* offsets, float* bunches) {
    if (!run) {
        set_last_error("null run");
        return MLGPU_ERR_ARGUMENT;
    }
    if (offsets) std::memcpy(offsets, run->offsets.data(), run->offsets.size() * 4);
    if (bunches) std::memcpy(bunches, run->bunches.data(), run->bunches.size() * 4);
    return MLGPU_OK;
}

int mlgpu_eventio_run_end(const mlgpu_eventio_run* run, float* raw) {
    if (!run || !raw) {
        set_last_error("null argument");
        return MLGPU_ERR_ARGUMENT;
    }
    if (!run->run_end_found) {
        set_last_error("eventio: the run end has not been read yet");
        return MLGPU_ERR_ARGUMENT;
    }
    std::memcpy(raw, run->run_end, sizeof(run->run_end));
    return MLGPU_OK;
}

}  // extern "C"
