"""Region queries on the device: BAI/CSI chunks read by the query server and
the overlap filter, against the oracle's index query and record filter
(crates/fritillaria-bam/src/io/reader/query.rs:104-152,
crates/fritillaria-csi/src/io/query.rs:66-82) on the samtools-indexed
fixture, on a synthetic three-reference BAM (BAI and CSI) and on a 200,000
record coordinate-sorted BAM."""
import numpy as np
import pytest

from conftest import golden_bytes
from helpers import bam_header, bgzf, build_record
from test_index import decode_file, multi_ref_bam, random_regions

from oracle import index as oi

pytestmark = pytest.mark.gpu


def check_regions(server, oracle_ix, fields, vo, regions, res, slot=0):
    sel = server.selection_to_host(slot)
    col_vo = server.slot_column_to_host(slot, "virtual_offsets")
    col_pos = server.slot_column_to_host(slot, "position")
    kept = 0
    for k, (rid, a, b) in enumerate(regions):
        r = res[k]
        assert r["status"] == 0, (k, rid, a, b, r)
        want, bad = oi.region_records(oracle_ix, fields, vo, rid, a, b)
        assert not bad
        chunks = oi.query(oracle_ix, rid, a, b)
        assert r["chunks"] == len(chunks)
        assert r["scanned"] == len(oi.query_records(chunks, vo))
        rows = sel[r["first"]:r["first"] + r["records"]]
        assert np.array_equal(col_vo[rows], vo[want]), (k, rid, a, b)
        assert np.array_equal(col_pos[rows], fields["pos"][want]), (k, rid, a, b)
        kept += len(want)
    return kept


def test_fixture_regions_in_one_batch(hip_ctx, oracle_mod):
    print("[gpu] regions: htslib_multiblock.bam with its BAI", flush=True)
    import fritillaria_amd as fa
    data = golden_bytes("htslib_multiblock.bam")
    u, rec, vo, vo_end, f = decode_file(oracle_mod, data)
    raw = golden_bytes("htslib_multiblock.bam.bai")
    ix, oix = fa.BinningIndex(raw), oi.read_index(raw)
    server = fa.QueryServer(hip_ctx, data)
    rng = np.random.default_rng(41)
    regions = random_regions(rng, 150, 420_000) + [(0, 1, 5_000_000), (0, None, None), (0, 6_000_000, None)]
    res = server.run_regions(ix, regions)
    kept = check_regions(server, oix, f, vo, regions, res)
    assert kept > 4000
    # one region alone gives the same rows as inside the batch
    one = server.run_regions(ix, [regions[3]])
    sel = server.selection_to_host(0)
    vo_col = server.slot_column_to_host(0, "virtual_offsets")
    want, _ = oi.region_records(oix, f, vo, *regions[3])
    assert np.array_equal(vo_col[sel[one[0]["first"]:one[0]["first"] + one[0]["records"]]], vo[want])
    server.close()


@pytest.mark.parametrize("kind", ["bai", "csi"])
def test_three_references_two_slots(hip_ctx, oracle_mod, kind):
    print(f"[gpu] regions: three references, {kind}, two slots", flush=True)
    import fritillaria_amd as fa
    data, refs = multi_ref_bam()
    u, rec, vo, vo_end, f = decode_file(oracle_mod, data)
    built = oi.build_index(oi.alignment_contexts(f, vo, vo_end), n_ref=len(refs), kind=kind)
    raw = oi.write_bai(built) if kind == "bai" else oi.write_csi(built)
    ix, oix = fa.BinningIndex(raw), oi.read_index(raw)
    server = fa.QueryServer(hip_ctx, data, slots=2)
    rng = np.random.default_rng(43)
    a_regions = random_regions(rng, 120, 400_000, n_ref=3)
    b_regions = random_regions(rng, 120, 400_000, n_ref=3)
    server.launch_regions(0, ix, a_regions)
    server.launch_regions(1, ix, b_regions)
    res_b = server.finish_regions(1)
    res_a = server.finish_regions(0)
    check_regions(server, oix, f, vo, a_regions, res_a, slot=0)
    check_regions(server, oix, f, vo, b_regions, res_b, slot=1)
    # an id the index lacks fails alone; the empty reference keeps nothing
    res = server.run_regions(ix, [(5, 1, 100), (1, 1, 400_000), (-1, None, None), a_regions[0]])
    assert res["status"].tolist()[:3] == [fa.IoError.kind, 0, fa.IoError.kind]
    assert res[1]["records"] == 0 and res[1]["chunks"] == 0
    check_regions(server, oix, f, vo, [a_regions[0]], res[3:], slot=0)
    server.close()


def test_a_bad_cigar_op_fails_its_region_only(hip_ctx, oracle_mod):
    print("[gpu] regions: a record failing the region check", flush=True)
    import fritillaria_amd as fa
    recs = [build_record(b"a", b"ACGT", pos=10, cigar=[(4 << 4) | 0], next_ref=-1, next_pos=-1, tlen=0),
            build_record(b"b", b"ACGT", pos=20, cigar=[(4 << 4) | 9], next_ref=-1, next_pos=-1, tlen=0),
            build_record(b"c", b"ACGT", pos=30, cigar=[(4 << 4) | 0], next_ref=-1, next_pos=-1, tlen=0)]
    data = bgzf(bam_header(text=b"@HD\tVN:1.6\tSO:coordinate\n") + b"".join(recs))
    u, rec, vo, vo_end, f = decode_file(oracle_mod, data)
    raw = oi.write_bai(oi.build_index(oi.alignment_contexts(f, vo, vo_end), n_ref=1))
    ix = fa.BinningIndex(raw)
    server = fa.QueryServer(hip_ctx, data)
    res = server.run_regions(ix, [(0, 1, 100), (0, None, None), (0, 11, 14)])
    # bounded: the second record's CIGAR is read and fails; unbounded: kept
    assert res[0]["status"] == fa.Malformed.kind and res[0]["error_position"] == 1
    assert res[1]["status"] == 0 and res[1]["records"] == 3
    assert res[2]["status"] == fa.Malformed.kind
    server.close()


@pytest.mark.parametrize("kind,min_shift,depth", [("bai", 14, 5), ("csi", 14, 5), ("csi", 12, 6)])
def test_sorted_200k_records(hip_ctx, oracle_mod, kind, min_shift, depth):
    print(f"[gpu] regions: 200,000 sorted records, {kind} min_shift {min_shift} depth {depth}", flush=True)
    import fritillaria_amd as fa
    from fritillaria_amd import synth
    s = synth.config_bam(1, n_records=200_000)
    u, rec, vo, vo_end, f = decode_file(oracle_mod, s.compressed)
    built = oi.build_index(oi.alignment_contexts(f, vo, vo_end), n_ref=1, min_shift=min_shift, depth=depth, kind=kind)
    raw = oi.write_bai(built) if kind == "bai" else oi.write_csi(built)
    ix, oix = fa.BinningIndex(raw), oi.read_index(raw)
    assert ix.format == kind and ix.binning == (min_shift, depth)
    server = fa.QueryServer(hip_ctx, s.compressed)
    rng = np.random.default_rng(47)
    regions = [(0, a, a + w) for a, w in zip(rng.integers(1, 1_000_000, 400).tolist(),
                                             rng.choice([1, 150, 2000, 50_000], 400).tolist())]
    regions += [(0, None, None)]
    res = server.run_regions(ix, regions)
    kept = check_regions(server, oix, f, vo, regions, res)
    assert res[-1]["records"] == 200_000 and kept > 200_000
    server.close()
