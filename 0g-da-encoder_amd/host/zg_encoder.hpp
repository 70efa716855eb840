// zg_encoder.hpp -- C++ host mirror of the reference encode API over include/zgr.h.
//
// Mirrors the types a caller of zg_encoder / amt uses on the encoding path:
//   RawData          crates/encoder/src/raw_data.rs:55-73      (TooLargeBlob)
//   EncoderParams    crates/amt/src/blob/encode.rs:21-133      (from_dir_mont, warmup)
//   EncodedBlob      crates/encoder/src/encoder/blob.rs:26-72  (build, get_row, commitment, roots)
//   EncodedSlice     crates/encoder/src/encoder/slice.rs        (row + AMT proof + Merkle proof)
//   EncoderService   grpc/src/service.rs:67-106                 (process_data reply fields)
// The compute runs in lib/libzgr.so on the GPU; this header holds no arithmetic.
// Errors follow the reference: an oversize payload throws TooLargeBlob, device and
// format failures throw std::runtime_error.
#pragma once

#include <array>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/zgr.h"

namespace zg {

using Bytes32 = std::array<uint8_t, 32>;

struct TooLargeBlob : std::runtime_error {  // EncoderError::TooLargeBlob
    size_t actual, expected_max;
    TooLargeBlob(size_t a, size_t m)
        : std::runtime_error("TooLargeBlob { actual: " + std::to_string(a) + ", expected_max: " + std::to_string(m) +
                             " }"),
          actual(a),
          expected_max(m) {}
};

inline void check(int rc, const char* what) {
    if (rc == ZGR_OK) return;
    throw std::runtime_error(std::string(what) + ": " + zgr_last_error() + " (code " + std::to_string(rc) + ")");
}

// Blob geometry (crates/encoder/src/constants.rs:14-28) as runtime values.
struct Geometry {
    int log_col = 10, log_row = 10, coset_n = 3;
    size_t n() const { return (size_t)1 << (log_col + log_row); }
    size_t row_n() const { return (size_t)1 << log_row; }
    size_t col_n() const { return (size_t)1 << log_col; }
    size_t row_encoded() const { return row_n() * coset_n; }
    size_t max_blob_size() const { return 31 * n(); }
    size_t max_raw_data_size() const { return max_blob_size() - 4; }
    size_t tree_pts() const { return ((size_t)1 << (log_row + 1)) - 2; }
    size_t merkle_nodes() const { return (size_t)coset_n * (((size_t)1 << (log_row + 1)) - 1); }
    size_t params_bytes() const { return (size_t)(log_row + 2) * n() * 64; }
    static Geometry production() { return Geometry{10, 10, 3}; }
    static Geometry test_setting() { return Geometry{5, 6, 3}; }
};

// RawData::try_from: only the size rule lives on the host; padding and packing run
// on the device (the 31-byte units and the u32 LE length footer).
struct RawData {
    const uint8_t* ptr = nullptr;
    size_t len = 0;
    static RawData try_from(const Geometry& g, const uint8_t* p, size_t n) {
        if (n > g.max_raw_data_size()) throw TooLargeBlob(n, g.max_raw_data_size());
        return RawData{p, n};
    }
};

// Prove parameters resident on one GPU for one geometry (EncoderParams<Bn254, COSET_N,
// LOG_COL, LOG_ROW>); owns the zgr context.
class EncoderParams {
  public:
    EncoderParams(const Geometry& g, int device = 0) : geo_(g) {
        zgr_ctx* c = nullptr;
        check(zgr_ctx_create(device, g.log_col, g.log_row, g.coset_n, &c), "zgr_ctx_create");
        ctx_.reset(c);
    }

    // insecure test setup from a known tau (PowerTau::setup_inner + AMTParams::from_pp)
    static EncoderParams from_tau(const Geometry& g, const uint8_t tau_le[32], int device = 0) {
        EncoderParams p(g, device);
        check(zgr_params_generate(p.ctx(), tau_le), "zgr_params_generate");
        return p;
    }

    // AMTParams::from_dir_mont (crates/amt/src/prove_params/generate.rs:27-65): read the
    // bamt files amt-prove-coset{c}-mont-{hash}-{h:02}-{n:02}.bin and keep the prove
    // bases resident.  The G2 sections (g2, high_g2, vanishes) belong to the verifier
    // and are skipped.  bamt layout: fast_serde_bn254.rs:20-88.
    static EncoderParams from_dir_mont(const Geometry& g, const std::string& dir,
                                       const std::string& type_hash = "5DWgDV", int device = 0) {
        EncoderParams p(g, device);
        const int depth = g.log_col + g.log_row, h = g.log_row;
        for (int c = 0; c < g.coset_n; c++) {
            char name[128];
            snprintf(name, sizeof(name), "amt-prove-coset%d-mont-%s-%02d-%02d.bin", c, type_hash.c_str(), h, depth);
            p.load_bamt(c, dir + "/" + name);
        }
        return p;
    }

    // read one bamt file and upload its prove bases (zgr_params_load_bamt parses the
    // layout and skips the G2 sections)
    void load_bamt(int coset, const std::string& path) {
        std::unique_ptr<FILE, int (*)(FILE*)> f(fopen(path.c_str(), "rb"), fclose);
        if (!f) throw std::runtime_error("cannot open " + path);
        if (fseek(f.get(), 0, SEEK_END)) throw std::runtime_error(path + ": seek");
        const long total = ftell(f.get());
        if (total < 0 || fseek(f.get(), 0, SEEK_SET)) throw std::runtime_error(path + ": seek");
        std::vector<uint8_t> image((size_t)total);
        if (fread(image.data(), 1, image.size(), f.get()) != image.size())
            throw std::runtime_error(path + ": truncated");
        check(zgr_params_load_bamt(ctx(), coset, image.data(), image.size()), ("zgr_params_load_bamt " + path).c_str());
    }

    void warmup() const {}  // bases are uploaded when loaded (EncoderParams::warmup)

    zgr_ctx* ctx() const { return ctx_.get(); }
    const Geometry& geometry() const { return geo_; }

  private:
    struct Del {
        void operator()(zgr_ctx* c) const { zgr_ctx_destroy(c); }
    };
    Geometry geo_;
    std::unique_ptr<zgr_ctx, Del> ctx_;
};

// ark-serialize 0.4 uncompressed G1 affine from canonical x || y (64 zero bytes =
// infinity): y carries the SW flags in its spare top bits, 0x40 in the last byte for
// infinity, 0x80 when y > -y (y > (p - 1) / 2).
inline std::array<uint8_t, 64> ark_g1_uncompressed(const std::array<uint8_t, 64>& p) {
    static const uint64_t half[4] = {0x9e10460b6c3e7ea3ull, 0xcbc0b548b438e546ull, 0xdc2822db40c0ac2eull,
                                     0x183227397098d014ull};  // (p - 1) / 2, little-endian limbs
    std::array<uint8_t, 64> r = p;
    bool zero = true;
    for (uint8_t b : p) zero = zero && b == 0;
    if (zero) {
        r[63] = 0x40;
        return r;
    }
    uint64_t y[4];
    memcpy(y, p.data() + 32, 32);
    for (int i = 3; i >= 0; i--) {
        if (y[i] == half[i]) continue;
        if (y[i] > half[i]) r[63] |= 0x80;
        break;
    }
    return r;
}

// One encoded row as EncodedSlice carries it (encoder/slice.rs, amt/slice.rs,
// merkle/slice.rs): index, row data, the AMT proof (commitment, proof) per level,
// the high commitment, the authoritative commitment, the Merkle proof.
struct EncodedSlice {
    size_t index = 0;
    size_t row_n = 0;  // rows per coset (BlobRow.index = index % row_n)
    std::vector<Bytes32> row;
    std::vector<std::pair<std::array<uint8_t, 64>, std::array<uint8_t, 64>>> amt_proof;
    std::array<uint8_t, 64> high_commitment{}, commitment{};
    std::vector<Bytes32> merkle_root;  // COSET_N roots
    std::vector<Bytes32> merkle_proof;
    Bytes32 row_root{};

    // serialize_uncompressed (grpc/src/service.rs:162-166) of EncodedSlice { index,
    // amt: EncodedSliceAMT { index, commitment, row: BlobRow { index, row, proof,
    // high_commitment } }, merkle: EncodedSliceMerkle { root, proof, leaf_index, leaf } }:
    // fields in declaration order, usize as u64 LE, Vec with a u64 LE length, arrays
    // without one, Fr as 32 B canonical LE.
    std::vector<uint8_t> serialize() const {
        std::vector<uint8_t> o;
        auto u64 = [&](uint64_t v) {
            for (int i = 0; i < 8; i++) o.push_back((uint8_t)(v >> (8 * i)));
        };
        auto put = [&](const uint8_t* p, size_t n) { o.insert(o.end(), p, p + n); };
        auto g1 = [&](const std::array<uint8_t, 64>& p) {
            auto a = ark_g1_uncompressed(p);
            put(a.data(), 64);
        };
        o.reserve(160 + 40 * row.size() + 136 * amt_proof.size() + 32 * (merkle_root.size() + merkle_proof.size()));
        u64(index);                     // EncodedSlice.index
        u64(index);                     // EncodedSliceAMT.index
        g1(commitment);
        u64(index % row_n);             // BlobRow.index
        u64(row.size());                // BlobRow.row
        for (const auto& e : row) put(e.data(), 32);
        u64(amt_proof.size());          // BlobRow.proof
        for (const auto& cp : amt_proof) {
            g1(cp.first);
            g1(cp.second);
        }
        g1(high_commitment);
        for (const auto& m : merkle_root) put(m.data(), 32);  // [Bytes32; COSET_N]
        u64(merkle_proof.size());
        for (const auto& m : merkle_proof) put(m.data(), 32);
        u64(index);                     // leaf_index
        put(row_root.data(), 32);       // leaf
        return o;
    }
};

// LightEncodedSlice (encoder/light_slice.rs:11-79): an EncodedSlice without its row.
struct LightEncodedSlice {
    size_t index = 0;
    std::array<uint8_t, 64> amt_commitment{}, amt_high_commitment{};
    std::vector<std::pair<std::array<uint8_t, 64>, std::array<uint8_t, 64>>> amt_proof;
    std::vector<Bytes32> merkle_root, merkle_proof;
    Bytes32 merkle_leaf{};

    // from_slice (light_slice.rs:59-79); the slice's AMT and Merkle indices are its
    // own index by construction (EncodedBlob::get_row)
    static LightEncodedSlice from_slice(const EncodedSlice& s) {
        LightEncodedSlice l;
        l.index = s.index;
        l.amt_commitment = s.commitment;
        l.amt_proof = s.amt_proof;
        l.amt_high_commitment = s.high_commitment;
        l.merkle_root = s.merkle_root;
        l.merkle_proof = s.merkle_proof;
        l.merkle_leaf = s.row_root;
        return l;
    }

    // into_slice (light_slice.rs:22-40): BlobRow.index = index % row_n
    EncodedSlice into_slice(std::vector<Bytes32> row, size_t row_n) const {
        EncodedSlice s;
        s.index = index;
        s.row_n = row_n;
        s.row = std::move(row);
        s.amt_proof = amt_proof;
        s.high_commitment = amt_high_commitment;
        s.commitment = amt_commitment;
        s.merkle_root = merkle_root;
        s.merkle_proof = merkle_proof;
        s.row_root = merkle_leaf;
        return s;
    }

    // serialize_uncompressed: index, amt_commitment, amt_proof, amt_high_commitment,
    // merkle_root, merkle_proof, merkle_leaf
    std::vector<uint8_t> serialize() const {
        std::vector<uint8_t> o;
        auto u64 = [&](uint64_t v) {
            for (int i = 0; i < 8; i++) o.push_back((uint8_t)(v >> (8 * i)));
        };
        auto put = [&](const uint8_t* p, size_t n) { o.insert(o.end(), p, p + n); };
        auto g1 = [&](const std::array<uint8_t, 64>& p) {
            auto a = ark_g1_uncompressed(p);
            put(a.data(), 64);
        };
        u64(index);
        g1(amt_commitment);
        u64(amt_proof.size());
        for (const auto& cp : amt_proof) {
            g1(cp.first);
            g1(cp.second);
        }
        g1(amt_high_commitment);
        for (const auto& m : merkle_root) put(m.data(), 32);
        u64(merkle_proof.size());
        for (const auto& m : merkle_proof) put(m.data(), 32);
        put(merkle_leaf.data(), 32);
        return o;
    }
};

class EncodedBlob {
  public:
    // EncodedBlob::build over RawData (RawBlob packing happens on the device)
    static EncodedBlob build(const RawData& raw, const EncoderParams& params) {
        EncodedBlob b(params.geometry());
        zgr_encoded o = b.outputs();
        check(zgr_encode(params.ctx(), raw.ptr, raw.len, &o), "zgr_encode");
        return b;
    }

    std::array<uint8_t, 64> get_affine_commitment() const { return point(commitment_.data()); }
    Bytes32 get_file_root() const {
        Bytes32 r;
        memcpy(r.data(), root_.data(), 32);
        return r;
    }
    std::vector<Bytes32> get_roots() const { return layer(0); }
    const std::vector<uint8_t>& get_data() const { return data_; }

    // EncodedBlob::get_row (encoder/blob.rs:44-49)
    EncodedSlice get_row(size_t index) const {
        const Geometry& g = geo_;
        if (index >= g.row_encoded()) throw std::out_of_range("row index");
        EncodedSlice s;
        s.index = index;
        s.row_n = g.row_n();
        const size_t coset = index / g.row_n(), local = index % g.row_n();
        for (size_t c = 0; c < g.col_n(); c++) {
            Bytes32 e;
            memcpy(e.data(), data_.data() + 32 * (coset * g.n() + local * g.col_n() + c), 32);
            s.row.push_back(e);
        }
        // HalfBlob::get_row + AllProofs::get_proof with reversed_index = bitrev(local)
        size_t rev = 0;
        for (int b = 0; b < g.log_row; b++) rev |= ((local >> b) & 1) << (g.log_row - 1 - b);
        for (int d = 0; d < g.log_row; d++) {
            size_t lv = rev >> (g.log_row - 1 - d);
            size_t at = coset * g.tree_pts() + ((size_t)2 << d) - 2 + (lv ^ 1);
            s.amt_proof.push_back({point(comm_tree_.data() + 64 * at), point(proof_tree_.data() + 64 * at)});
        }
        s.high_commitment = point(high_.data() + 64 * coset);
        s.commitment = get_affine_commitment();
        // EncodedBlobMerkle::get_row (merkle/blob.rs:44-57)
        s.merkle_root = layer(0);
        for (int d = g.log_row; d >= 1; d--) s.merkle_proof.push_back(layer(d)[(index >> (g.log_row - d)) ^ 1]);
        s.row_root = layer(g.log_row)[index];
        return s;
    }

    // raw output buffers (zgr_encoded layout), for callers that copy them out
    const std::vector<uint8_t>& comm_tree() const { return comm_tree_; }
    const std::vector<uint8_t>& proof_tree() const { return proof_tree_; }
    const std::vector<uint8_t>& high() const { return high_; }
    const std::vector<uint8_t>& merkle() const { return merkle_; }

  private:
    explicit EncodedBlob(const Geometry& g)
        : geo_(g),
          data_((size_t)g.coset_n * g.n() * 32),
          commitment_(64),
          comm_tree_((size_t)g.coset_n * g.tree_pts() * 64),
          proof_tree_((size_t)g.coset_n * g.tree_pts() * 64),
          high_((size_t)g.coset_n * 64),
          merkle_(g.merkle_nodes() * 32),
          root_(32) {}

    zgr_encoded outputs() {
        return zgr_encoded{data_.data(), commitment_.data(), comm_tree_.data(), proof_tree_.data(),
                           high_.data(),  merkle_.data(),     root_.data()};
    }
    static std::array<uint8_t, 64> point(const uint8_t* p) {
        std::array<uint8_t, 64> r;
        memcpy(r.data(), p, 64);
        return r;
    }
    std::vector<Bytes32> layer(int d) const {  // EncodedBlobMerkle.tree[d]
        size_t off = 0;
        for (int k = 0; k < d; k++) off += (size_t)geo_.coset_n << k;
        std::vector<Bytes32> v((size_t)geo_.coset_n << d);
        for (size_t i = 0; i < v.size(); i++) memcpy(v[i].data(), merkle_.data() + 32 * (off + i), 32);
        return v;
    }

    Geometry geo_;
    std::vector<uint8_t> data_, commitment_, comm_tree_, proof_tree_, high_, merkle_, root_;
};

// EncodeBlobReply (grpc/proto/encoder.proto:16-22) as EncoderService::process_data
// fills it (grpc/src/service.rs:67-106).
struct EncodeBlobReply {
    uint32_t version = 0;
    std::vector<uint8_t> erasure_commitment;  // x || y, 32 B canonical LE each (service.rs:77-83)
    std::vector<uint8_t> storage_root;        // file root (service.rs:84)
    std::vector<uint8_t> encoded_data;        // 32 B per encoded element when require_data (service.rs:85-91)
    std::vector<std::vector<uint8_t>> encoded_slice;  // serialized get_row(i) for every row (service.rs:93-95)

    // proto3 encoding of the message: fields with default values are not written,
    // every element of the repeated field is
    std::vector<uint8_t> encode() const {
        std::vector<uint8_t> o;
        size_t total = 16 + erasure_commitment.size() + storage_root.size() + encoded_data.size();
        for (const auto& s : encoded_slice) total += 16 + s.size();
        o.reserve(total);
        if (version) {
            proto_varint(o, 1 << 3);
            proto_varint(o, version);
        }
        proto_bytes(o, 2, erasure_commitment, false);
        proto_bytes(o, 3, storage_root, false);
        proto_bytes(o, 4, encoded_data, false);
        for (const auto& s : encoded_slice) proto_bytes(o, 5, s, true);
        return o;
    }

    static void proto_varint(std::vector<uint8_t>& o, uint64_t v) {
        while (v >= 0x80) {
            o.push_back((uint8_t)(v | 0x80));
            v >>= 7;
        }
        o.push_back((uint8_t)v);
    }

    static void proto_bytes(std::vector<uint8_t>& o, int field, const std::vector<uint8_t>& b, bool always) {
        if (b.empty() && !always) return;
        proto_varint(o, ((uint64_t)field << 3) | 2);
        proto_varint(o, b.size());
        o.insert(o.end(), b.begin(), b.end());
    }
};

// EncodeBlobRequest (grpc/proto/encoder.proto:12-15): bytes data = 1; bool require_data = 2.
struct EncodeBlobRequest {
    std::vector<uint8_t> data;
    bool require_data = false;

    // proto3 decoding: unknown fields are skipped, a repeated scalar keeps its last
    // value; malformed input throws std::invalid_argument
    static EncodeBlobRequest decode(const uint8_t* p, size_t n) {
        EncodeBlobRequest r;
        size_t i = 0;
        auto varint = [&]() -> uint64_t {
            uint64_t v = 0;
            for (int shift = 0; shift < 70; shift += 7) {
                if (i >= n) throw std::invalid_argument("truncated varint");
                uint8_t b = p[i++];
                v |= (uint64_t)(b & 0x7f) << shift;
                if (!(b & 0x80)) return v;
            }
            throw std::invalid_argument("varint too long");
        };
        auto skip = [&](size_t k) {
            if (k > n - i) throw std::invalid_argument("truncated field");
            i += k;
        };
        while (i < n) {
            const uint64_t key = varint();
            const uint64_t field = key >> 3, wt = key & 7;
            if (field == 0) throw std::invalid_argument("field number 0");
            if (wt == 0) {
                const uint64_t v = varint();
                if (field == 2) r.require_data = v != 0;
                else if (field == 1) throw std::invalid_argument("field 1 is not a varint");
            } else if (wt == 2) {
                const uint64_t len = varint();
                if (len > n - i) throw std::invalid_argument("truncated length-delimited field");
                if (field == 1) r.data.assign(p + i, p + i + len);
                else if (field == 2) throw std::invalid_argument("field 2 is not a varint");
                i += len;
            } else if (wt == 1 || wt == 5) {
                if (field == 1 || field == 2) throw std::invalid_argument("fixed-width known field");
                skip(wt == 1 ? 8 : 4);
            } else {
                throw std::invalid_argument("unsupported wire type");
            }
        }
        return r;
    }
};

class EncoderService {
  public:
    explicit EncoderService(EncoderParams params) : params_(std::move(params)) {}

    EncodeBlobReply process_data(const uint8_t* data, size_t len, bool require_data) const {
        RawData raw = RawData::try_from(params_.geometry(), data, len);
        EncodedBlob blob = EncodedBlob::build(raw, params_);
        EncodeBlobReply r;
        auto c = blob.get_affine_commitment();
        r.erasure_commitment.assign(c.begin(), c.end());
        auto root = blob.get_file_root();
        r.storage_root.assign(root.begin(), root.end());
        if (require_data) r.encoded_data = blob.get_data();
        r.encoded_slice.resize(params_.geometry().row_encoded());
        for (size_t i = 0; i < r.encoded_slice.size(); i++) r.encoded_slice[i] = blob.get_row(i).serialize();
        return r;
    }

    const EncoderParams& params() const { return params_; }

  private:
    EncoderParams params_;
};

// recover_from_da_slice (crates/recovery/src/lib.rs:15-43) on the device: the raw data
// of a blob from any 2^log_row of its rows, keyed by row index (EncodedSlice::merkle_row
// bytes).  Rows of the wrong length are dropped, as the reference does.  Throws
// RecoveryError with the reference's Err text, RecoveryPanic where it panics.
struct RecoveryError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct RecoveryPanic : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline std::vector<uint8_t> recover_from_da_slice(const EncoderParams& params,
                                                  const std::map<size_t, std::vector<uint8_t>>& slices) {
    const Geometry& g = params.geometry();
    const size_t width = 32 * g.col_n();
    std::vector<uint32_t> ids;
    std::vector<uint8_t> rows;
    for (const auto& kv : slices) {
        if (kv.second.size() != width) continue;
        if (kv.first > UINT32_MAX) throw RecoveryError("Cannot recover data: RowIdOverflow");
        ids.push_back((uint32_t)kv.first);
        rows.insert(rows.end(), kv.second.begin(), kv.second.end());
    }
    std::vector<uint8_t> out(31 * g.n() - 4);
    size_t len = 0;
    const int rc = zgr_recover(params.ctx(), ids.data(), rows.data(), ids.size(), out.data(), out.size(), &len);
    if (rc != ZGR_OK) {
        const std::string msg = zgr_last_error();
        if (rc == ZGR_ERR_INCONSISTENT || msg.find("panics") != std::string::npos) throw RecoveryPanic(msg);
        if (rc == ZGR_ERR_INVALID_ARG) throw RecoveryError(msg);
        check(rc, "zgr_recover");
    }
    out.resize(len);
    return out;
}

}  // namespace zg
