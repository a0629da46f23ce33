"""The profile summaries' bookkeeping (CPU only): how a kernel trace is cut
into update batches, and how HBM bytes follow from request-size counters."""
import profile_summary as ps


def d(name):
    return (name, None)


def test_batches_start_at_either_first_kernel():
    trace = [d("k_fill_u64"), d("k_markers"),
             # a bulk load: batch start, then the large-batch radix sort
             d("k_batch_begin"), d("k_pack_group"), d("k_rs_hist"), d("k_rs_scatter"), d("k_pull"),
             # a lone PMA's small batch: the split kernel carries the batch start
             d("k_ss_split"), d("k_ss_bucket"), d("k_locate"), d("k_compact"), d("k_density"),
             # a batch on another path: k_batch_begin, then the same sort
             d("k_batch_begin"), d("k_ss_split"), d("k_ss_bucket"), d("k_locate"),
             # a batch of the one-sweep sort: with the radix sorts it counts as
             # a bulk load (update batches of the headline sort in two kernels)
             d("k_batch_begin"), d("k_pack_count"), d("k_os_pass"), d("k_locate"),
             # two carried batches in a row
             d("k_ss_split"), d("k_ss_bucket"), d("k_rebalance_seg"),
             d("k_ss_split"), d("k_ss_bucket")]
    upd, bulk, pre = ps.batches(trace)
    assert [k for k, _ in pre] == ["k_fill_u64", "k_markers"]
    assert [b[1][0] for b in bulk] == ["k_pack_group", "k_pack_count"]
    firsts = [[k for k, _ in b][:2] for b in upd]
    assert firsts == [["k_ss_split", "k_ss_bucket"], ["k_batch_begin", "k_ss_split"],
                      ["k_ss_split", "k_ss_bucket"], ["k_ss_split", "k_ss_bucket"]]
    # every dispatch is in exactly one place
    assert len(pre) + sum(len(b) for b in upd) + sum(len(b) for b in bulk) == len(trace)


def test_request_bytes():
    c = {"TCC_EA0_RDREQ_32B": 3, "TCC_EA0_RDREQ_64B": 5, "TCC_EA0_RDREQ_128B": 7,
         "TCC_EA0_WRREQ_64B": 11, "TCC_EA0_WRREQ": 13}
    rd, wr = ps.req_bytes(c)
    assert rd == 32 * 3 + 64 * 5 + 128 * 7
    assert wr == 64 * 11 + 32 * 2  # the write requests that are not 64 B are 32 B
    assert ps.req_bytes({}) == (0, 0)


def test_short_kernel_names():
    assert ps.short("void pma::k_ss_split<true>(pma::BatchSeg const*, unsigned int)") == "k_ss_split"
    assert ps.short("pma::k_locate<false>(pma::Group)") == "k_locate"
    assert ps.short("__amd_rocclr_copyBuffer") == "__amd_rocclr_copyBuffer"
