// zg_encode_file -- encode raw blob files through the C++ host mirror (zg_encoder.hpp).
//
//   zg_encode_file LOG_COL LOG_ROW COSET_N PARAMS OUTDIR RAW... [--rows I,J,...] [--pb]
//                  [--recover I,J,...]
//     PARAMS  tau:<64 hex digits, 32 B little-endian>   insecure setup from a known tau
//             dir:<path>                                bamt prove-param files
//                                                       (amt-prove-coset{c}-mont-5DWgDV-{h}-{n}.bin)
//   For RAW number k it writes OUTDIR/k.reply = erasure_commitment (64) || storage_root (32)
//   || encoded_data (EncoderService::process_data with require_data), OUTDIR/k.slices = the
//   reply's encoded_slice bytes of every row, concatenated in row order, and, with --rows,
//   OUTDIR/k.rows = for each index: row || per level (commitment, proof) || high_commitment
//   || commitment || merkle roots || merkle proof (leaf level first) || row root
//   (EncodedBlob::get_row), and OUTDIR/k.light = the LightEncodedSlice bytes of those rows.
//   OUTDIR/k.pb is the reply as protobuf (the EncodeBlob response message).  With --pb
//   each RAW file is an EncodeBlobRequest message: its data is encoded and its
//   require_data decides whether k.pb carries encoded_data (k.reply always does).
//   With --recover, OUTDIR/k.recovered = recover_from_da_slice over the listed rows of
//   the blob's own encoding (the reference's Err / panic text goes to stderr, exit 5).
//   Exit codes: 0 ok, 2 usage, 3 TooLargeBlob, 4 other error, 5 recovery error.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <sstream>

#include "zg_encoder.hpp"

namespace {

std::vector<uint8_t> read_file(const std::string& p) {
    std::ifstream f(p, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + p);
    return std::vector<uint8_t>(std::istreambuf_iterator<char>(f), {});
}

void write_file(const std::string& p, const std::vector<uint8_t>& b) {
    std::ofstream f(p, std::ios::binary);
    f.write(reinterpret_cast<const char*>(b.data()), (std::streamsize)b.size());
    if (!f) throw std::runtime_error("cannot write " + p);
}

template <class C>
void put(std::vector<uint8_t>& o, const C& c) {
    o.insert(o.end(), c.begin(), c.end());
}

bool parse_tau(const std::string& hex, uint8_t tau[32]) {
    if (hex.size() != 64) return false;
    for (int i = 0; i < 32; i++) {
        char* end = nullptr;
        std::string byte = hex.substr(2 * i, 2);
        long v = strtol(byte.c_str(), &end, 16);
        if (*end) return false;
        tau[i] = (uint8_t)v;
    }
    return true;
}

int usage() {
    fprintf(stderr, "usage: zg_encode_file LOG_COL LOG_ROW COSET_N tau:<hex>|dir:<path> OUTDIR RAW... [--rows I,J,...] [--pb] [--recover I,J,...]\n");
    return 2;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 7) return usage();
    zg::Geometry g{atoi(argv[1]), atoi(argv[2]), atoi(argv[3])};
    const std::string params = argv[4], outdir = argv[5];
    std::vector<std::string> raws;
    std::vector<size_t> rows, recover_rows;
    bool pb = false;
    for (int i = 6; i < argc; i++) {
        std::string a = argv[i];
        if (a == "--rows" && i + 1 < argc) {
            std::stringstream ss(argv[++i]);
            std::string tok;
            while (std::getline(ss, tok, ',')) rows.push_back(std::stoul(tok));
        } else if (a == "--recover" && i + 1 < argc) {
            std::stringstream ss(argv[++i]);
            std::string tok;
            while (std::getline(ss, tok, ',')) recover_rows.push_back(std::stoul(tok));
        } else if (a == "--pb") {
            pb = true;
        } else {
            raws.push_back(a);
        }
    }
    if (raws.empty()) return usage();
    try {
        zg::EncoderParams p = [&]() {
            if (params.rfind("tau:", 0) == 0) {
                uint8_t tau[32];
                if (!parse_tau(params.substr(4), tau)) throw std::invalid_argument("tau: needs 64 hex digits");
                return zg::EncoderParams::from_tau(g, tau);
            }
            if (params.rfind("dir:", 0) == 0) return zg::EncoderParams::from_dir_mont(g, params.substr(4));
            throw std::invalid_argument("PARAMS must be tau:<hex> or dir:<path>");
        }();
        p.warmup();
        zg::EncoderService svc(std::move(p));
        for (size_t k = 0; k < raws.size(); k++) {
            const std::vector<uint8_t> file = read_file(raws[k]);
            zg::EncodeBlobRequest req;
            if (pb) {
                req = zg::EncodeBlobRequest::decode(file.data(), file.size());
            } else {
                req.data = file;
                req.require_data = true;
            }
            const std::vector<uint8_t>& raw = req.data;
            const std::string stem = outdir + "/" + std::to_string(k);
            zg::EncodeBlobReply r = svc.process_data(raw.data(), raw.size(), true);
            if (!req.require_data) {
                zg::EncodeBlobReply lean = r;
                lean.encoded_data.clear();
                write_file(stem + ".pb", lean.encode());
            } else {
                write_file(stem + ".pb", r.encode());
            }
            std::vector<uint8_t> reply;
            put(reply, r.erasure_commitment);
            put(reply, r.storage_root);
            put(reply, r.encoded_data);
            write_file(stem + ".reply", reply);
            std::vector<uint8_t> slices;  // every encoded_slice, in row order
            for (const auto& sl : r.encoded_slice) put(slices, sl);
            write_file(stem + ".slices", slices);
            if (!recover_rows.empty()) {
                const size_t w = 32 * g.col_n();
                std::map<size_t, std::vector<uint8_t>> lines;
                for (size_t idx : recover_rows) {
                    const size_t src = idx < g.row_encoded() ? idx : 0;  // ids past the blob: row 0's bytes
                    lines[idx].assign(r.encoded_data.begin() + src * w, r.encoded_data.begin() + (src + 1) * w);
                }
                try {
                    write_file(stem + ".recovered", zg::recover_from_da_slice(svc.params(), lines));
                } catch (const zg::RecoveryError& e) {
                    fprintf(stderr, "recovery error: %s\n", e.what());
                    return 5;
                } catch (const zg::RecoveryPanic& e) {
                    fprintf(stderr, "recovery panic: %s\n", e.what());
                    return 5;
                }
            }
            if (rows.empty()) continue;
            zg::EncodedBlob b = zg::EncodedBlob::build(zg::RawData::try_from(g, raw.data(), raw.size()), svc.params());
            std::vector<uint8_t> out, light;
            for (size_t idx : rows) {
                zg::EncodedSlice s = b.get_row(idx);
                put(light, zg::LightEncodedSlice::from_slice(s).serialize());
                for (const auto& e : s.row) put(out, e);
                for (const auto& cp : s.amt_proof) {
                    put(out, cp.first);
                    put(out, cp.second);
                }
                put(out, s.high_commitment);
                put(out, s.commitment);
                for (const auto& m : s.merkle_root) put(out, m);
                for (const auto& m : s.merkle_proof) put(out, m);
                put(out, s.row_root);
            }
            write_file(stem + ".rows", out);
            write_file(stem + ".light", light);
        }
    } catch (const zg::TooLargeBlob& e) {
        fprintf(stderr, "%s\n", e.what());
        return 3;
    } catch (const std::exception& e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 4;
    }
    return 0;
}
