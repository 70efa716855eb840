"""Ceremony import: powers of tau from a perpetual-powers-of-tau file, the host side of
crates/ppot2ark/src/main.rs (from_ppot_file_inner :45-119, from_ppot_file :121-135,
from_ppot_file_ldt :137-164, load_save_power_tau :166-191).

The reference reads the files through the powersoftau crate of kobigurk/phase2-bn254
(rev dd6b966; an absent dependency): a challenge holds the BatchedAccumulator with
uncompressed points, a response the same with compressed points (then the
contributor's public key), each behind a 64-byte BLAKE2b hash.  Sections, in order:
tau_powers_g1 (2 * 2^power - 1 points), tau_powers_g2, alpha_tau_powers_g1 and
beta_tau_powers_g1 (2^power each), beta_g2 (one point).  A chunk read
(BatchedAccumulator::read_chunk, CheckForCorrectness::Yes) decodes and checks the
chunk's range of every section and beta_g2; here the file is memory-mapped, each
section's range is handed to zgr_ppot_decode, and the decoding, square roots and
subgroup checks run on the device.

Points come back in the library's memory form: canonical affine, G1 x || y (64 B),
G2 x.c0 || x.c1 || y.c0 || y.c1 (128 B), little-endian.
"""
import enum
import os
from typing import NamedTuple

import numpy as np

from . import (BN254_HASH, FILE_PTAU, PPOT_G1_COMPRESSED, PPOT_G1_UNCOMPRESSED, PPOT_G2_COMPRESSED,
               PPOT_G2_UNCOMPRESSED, ZgrError)

HASH_SIZE = 64  # BLAKE2b-512 of the previous file, at the start of every file


class PpotError(RuntimeError):
    """The reference's Err(String) of from_ppot_file and friends."""


class InputType(enum.Enum):
    """crates/ppot2ark/src/main.rs:27-43."""
    Challenge = "challenge"
    Response = "response"

    def file_name(self, degree):
        return "%s_%d" % (self.value, degree)


class ElementType(enum.IntEnum):
    TauG1 = 0
    TauG2 = 1
    AlphaG1 = 2
    BetaG1 = 3
    BetaG2 = 4


class CeremonyParams:
    """CeremonyParams::<Bn256>::new(circuit_power, batch_size) of powersoftau: the section
    lengths of an accumulator of 2^circuit_power powers."""

    def __init__(self, circuit_power, batch_size):
        self.size = circuit_power
        self.batch_size = batch_size
        self.powers_length = 1 << circuit_power
        self.powers_g1_length = (self.powers_length << 1) - 1
        self.hash_size = HASH_SIZE

    @staticmethod
    def point_size(element, compressed):
        g2 = element in (ElementType.TauG2, ElementType.BetaG2)
        return (64 if g2 else 32) * (1 if compressed else 2)

    def section_length(self, element):
        if element == ElementType.TauG1:
            return self.powers_g1_length
        if element == ElementType.BetaG2:
            return 1
        return self.powers_length

    def position(self, index, element, compressed):
        """Byte offset of element `index` of a section (calculate_mmap_position)."""
        g1 = self.point_size(ElementType.TauG1, compressed)
        g2 = self.point_size(ElementType.TauG2, compressed)
        base = {
            ElementType.TauG1: 0,
            ElementType.TauG2: g1 * self.powers_g1_length,
            ElementType.AlphaG1: g1 * self.powers_g1_length + g2 * self.powers_length,
            ElementType.BetaG1: g1 * self.powers_g1_length + g2 * self.powers_length + g1 * self.powers_length,
            ElementType.BetaG2: g1 * self.powers_g1_length + g2 * self.powers_length + 2 * g1 * self.powers_length,
        }[ElementType(element)]
        return self.hash_size + base + index * self.point_size(element, compressed)

    def file_length(self, compressed):
        """Bytes up to the end of beta_g2 (a response then carries the public key)."""
        return self.position(0, ElementType.BetaG2, compressed) + self.point_size(ElementType.BetaG2, compressed)


def _kind(element, compressed):
    g2 = element in (ElementType.TauG2, ElementType.BetaG2)
    if g2:
        return PPOT_G2_COMPRESSED if compressed else PPOT_G2_UNCOMPRESSED
    return PPOT_G1_COMPRESSED if compressed else PPOT_G1_UNCOMPRESSED


class PowerTauLight(NamedTuple):
    """crates/ppot2ark/src/main.rs:23: the G1 and G2 powers read."""
    g1: np.ndarray  # (n, 64)
    g2: np.ndarray  # (n, 128)


class PowerTau(NamedTuple):
    """amt::PowerTau (crates/amt/src/power_tau.rs:20-26), canonical affine arrays."""
    g1pp: np.ndarray
    g2pp: np.ndarray
    high_g1pp: np.ndarray
    high_g2: np.ndarray  # (128,)


def read_points_chunk(ctx, input_map, params, start, size, element, compressed):
    """BatchedAccumulator::read_points_chunk: decode and check elements [start, start +
    size) of one section on the device.  A range running past the section gives no
    points (the crate returns an empty vector there); a file too short for the range is
    an error."""
    if start + size > params.section_length(element):
        return None
    per = params.point_size(element, compressed)
    pos = params.position(start, element, compressed)
    if pos + size * per > input_map.size:
        raise PpotError("must read point data from file")
    try:
        return ctx.ppot_decode(_kind(element, compressed), input_map[pos:pos + size * per], size)
    except ZgrError as e:
        raise PpotError("failed to read chunk, detail: %s[%d]: %s"
                        % (element.name, start + getattr(e, "first_bad", 0), e.msg)) from None


def read_chunk(ctx, input_map, params, start, size, compressed):
    """BatchedAccumulator::read_chunk: every section's range (checked), -> (tau_g1, tau_g2)."""
    tau_g1 = read_points_chunk(ctx, input_map, params, start, size, ElementType.TauG1, compressed)
    tau_g2 = read_points_chunk(ctx, input_map, params, start, size, ElementType.TauG2, compressed)
    read_points_chunk(ctx, input_map, params, start, size, ElementType.AlphaG1, compressed)
    read_points_chunk(ctx, input_map, params, start, size, ElementType.BetaG1, compressed)
    read_points_chunk(ctx, input_map, params, 0, 1, ElementType.BetaG2, compressed)
    return tau_g1, tau_g2


def from_ppot_file_inner(ctx, input_path, input_type, file_size, read_from, read_size_pow, chunk_size_pow,
                         parameters):
    """crates/ppot2ark/src/main.rs:45-119."""
    read_size = 1 << read_size_pow
    chunk_size = 1 << chunk_size_pow
    if read_from + read_size > (1 << file_size):
        raise PpotError("too long to read")
    input_filename = "%s/%s" % (input_path, input_type.file_name(file_size))
    try:
        input_map = np.memmap(input_filename, dtype=np.uint8, mode="r")
    except (OSError, ValueError) as e:
        raise PpotError("Cannot open %s: %r" % (input_filename, e)) from None
    compressed = input_type == InputType.Response
    g1, g2 = [], []
    offset, remaining = read_from, read_size
    while remaining > 0:
        current = min(chunk_size, remaining)
        tau_g1, tau_g2 = read_chunk(ctx, input_map, parameters, offset, current, compressed)
        if tau_g1 is not None:
            g1.append(tau_g1)
        if tau_g2 is not None:
            g2.append(tau_g2)
        offset += current
        remaining -= current
    del input_map
    cat = lambda parts, w: np.concatenate(parts) if parts else np.zeros((0, w), dtype=np.uint8)
    return PowerTauLight(cat(g1, 64), cat(g2, 128))


def from_ppot_file(ctx, input_path, input_type, file_size_pow, read_from, read_size_pow, chunk_size_pow):
    """crates/ppot2ark/src/main.rs:121-135: points [read_from, read_from + 2^read_size_pow)
    of tau_powers_g1 and tau_powers_g2 of <input_path>/<challenge|response>_<file_size_pow>."""
    params = CeremonyParams(file_size_pow, file_size_pow)
    return from_ppot_file_inner(ctx, input_path, input_type, file_size_pow, read_from, read_size_pow,
                                chunk_size_pow, params)


def from_ppot_file_ldt(ctx, input_path, input_type, file_size_pow, read_size_pow, high_read_size_pow,
                       chunk_size_pow):
    """crates/ppot2ark/src/main.rs:137-164: g1pp / g2pp = the first 2^read_size_pow powers,
    high_g1pp = the powers from H = 2^high_read_size_pow - 2^read_size_pow on, and
    high_g2 = [tau^H] G2."""
    g1pp, g2pp = from_ppot_file(ctx, input_path, input_type, file_size_pow, 0, read_size_pow, chunk_size_pow)
    high_read_from = (1 << high_read_size_pow) - (1 << read_size_pow)
    high_g1pp, high_g2pp = from_ppot_file(ctx, input_path, input_type, file_size_pow, high_read_from,
                                          read_size_pow, chunk_size_pow)
    return PowerTau(g1pp, g2pp, high_g1pp, high_g2pp[0].copy())


def ptau_file_name(depth, mont=False):
    """crates/amt/src/utils.rs:31-34 for Bn254."""
    return "power-tau%s-%s-%02d.bin" % ("-mont" if mont else "", BN254_HASH, depth)


def load_save_power_tau(ctx, input_path, input_type, file_size_pow, target_size_pow, high_read_size_pow,
                        chunk_size_pow, dir):
    """crates/ppot2ark/src/main.rs:166-191: read the powers and write the compressed
    PowerTau file <dir>/power-tau-5DWgDV-<target>.bin.  The powers stay loaded in ctx
    (zgr_power_tau_load), whose domain must be 2^target_size_pow, ready for
    ctx.params_from_power_tau.  Returns the file's path."""
    if ctx.geo.n != 1 << target_size_pow:
        raise PpotError("the context's domain holds 2^%d points, not 2^%d" % (ctx.geo.n.bit_length() - 1,
                                                                            target_size_pow))
    pt = from_ppot_file_ldt(ctx, input_path, input_type, file_size_pow, target_size_pow, high_read_size_pow,
                            chunk_size_pow)
    ctx.power_tau_load(pt.g1pp, pt.g2pp, pt.high_g1pp, pt.high_g2)
    path = os.path.join(dir, ptau_file_name(target_size_pow, False))
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "wb") as f:
        f.write(ctx.write_params_file(FILE_PTAU))
    return path
