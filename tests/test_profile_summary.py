"""The profile summaries' selection of the launches after the launch tuner's
sweep: each functor's last variant, from the dispatch after the last one of
any other variant of the same functor."""
from disco_amd.profile_summary import after_tuning, functor_of, timer_name

RR = "void k_run<dh::KGrad<8, 1> >(dh::KGrad<8, 1>, long)"
XCD = "void k_run_xcd<dh::KGrad<8, 1> >(dh::KGrad<8, 1>, long)"
STRIP = "void k_run_xcd<dh::Grouped<dh::KGrad<8, 1> > >(dh::Grouped<dh::KGrad<8, 1> >, long)"
OTHER = "void k_run<dh::KUpdate<8, dh::MHD<8> > >(dh::KUpdate<8, dh::MHD<8> >, long)"
COOP = "k_flips(dh::Mesh, int, dh::Faces, signed char const*, double*, double*)"


def test_functor_of_every_launch_variant():
    assert functor_of(RR) == functor_of(XCD) == functor_of(STRIP) == "dh::KGrad<8, 1>"
    assert functor_of(OTHER) == "dh::KUpdate<8, dh::MHD<8> >"
    assert functor_of(COOP) == COOP
    assert timer_name(STRIP) == "grad_r" and timer_name(COOP) == "flips"


def test_after_tuning_keeps_the_last_variant_after_the_sweep():
    # the sweep: variants in turn (dispatches 1-6), then the kept variant
    # (xcd) for the warmup and timed steps; another kernel throughout
    rows = [(RR, 1, 10.0), (XCD, 2, 8.0), (STRIP, 3, 9.0), (RR, 4, 10.0), (XCD, 5, 8.0),
            (STRIP, 6, 9.0), (XCD, 7, 7.0), (XCD, 8, 7.5), (OTHER, 2, 1.0), (OTHER, 9, 2.0),
            (COOP, 3, 0.5)]
    kept = after_tuning(rows)
    assert sorted(r[1] for r in kept if functor_of(r[0]) == "dh::KGrad<8, 1>") == [7, 8]
    assert all(r[0] == XCD for r in kept if functor_of(r[0]) == "dh::KGrad<8, 1>")
    # a kernel with one variant keeps every dispatch
    assert sorted(r[1] for r in kept if r[0] == OTHER) == [2, 9]
    assert [r[1] for r in kept if r[0] == COOP] == [3]


def test_after_tuning_when_the_kept_variant_was_swept_last():
    # the kept variant also ran last in the sweep: its sweep dispatch after
    # the other variants' last one belongs to the same variant and stays
    rows = [(RR, 1, 10.0), (XCD, 2, 8.0), (RR, 3, 10.0), (XCD, 4, 8.0), (XCD, 5, 7.0)]
    assert [r[1] for r in after_tuning(rows)] == [4, 5]
