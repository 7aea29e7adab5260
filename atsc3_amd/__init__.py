"""atsc3_amd -- MI355X-native BICM demapper and LDPC min-sum decoder.

Drop-in for the demap and min-sum calls of the Felbs/atsc3 ATSC 3.0
receiver.  Hand-written HIP kernels (gfx950) in ``csrc/`` behind the C ABI of
``include/atsc3_fec.h``; thin Python wrappers with the upstream signatures:

  atsc3_amd.ldpc   min_sum_decode_batch, GpuMinSum, ladder_decode, decode_dev
  atsc3_amd.demap  demap_llr, demap_batch, Demapper
  atsc3_amd.codes  A/322 LDPC codes (H, encoder, QC tables)
  atsc3_amd.bicm   constellations and bit-interleaver maps
"""
__all__ = ["codes", "bicm", "synth"]
