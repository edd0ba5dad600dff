"""Parameter sets shared by the tests (small versions of the benchmark configs)."""
import nanofiller_amd as nf


def cfg1(precision=nf.NF_PREC_FP64, **kw):
    """20,000 rods, l = 1 um, a = 5 nm, 5 um periodic box, seed 1, 1 um cells."""
    base = dict(box=(5, 5, 5), n_rods=20000, len=1.0, radius=0.005, seed=1, cell_size=1.0,
                contact_gap=0.002, precision=precision)
    base.update(kw)
    return nf.normalize(nf.default_params(**base))


def small_iso(precision=nf.NF_PREC_FP64, **kw):
    base = dict(box=(3, 3, 3), n_rods=3000, len=1.0, radius=0.005, seed=7, contact_gap=0.002,
                precision=precision)
    base.update(kw)
    return nf.normalize(nf.default_params(**base))


def small_cfg2(precision=nf.NF_PREC_FP32_RECHECK, **kw):
    """cfg2 laws at 1/64 volume: log-normal lengths, 8 agglomerates, vMF(+z, 5)."""
    base = dict(box=(5, 5, 5), n_rods=7800, len=1.0, len_law=nf.NF_LEN_LOGNORMAL, len_sigma=0.3,
                len_min=0.3, len_max=3.0, radius=0.005, seed=2, den_law=nf.NF_DEN_AGGLOM, n_agg=8,
                agg_sigma=1.5, agg_peak=3.0, ort_law=nf.NF_ORT_VMF, kappa=5.0, axis=(0, 0, 1),
                contact_gap=0.002, precision=precision)
    base.update(kw)
    return nf.normalize(nf.default_params(**base))


def cfg2(precision=nf.NF_PREC_FP32_RECHECK, **kw):
    """500,000 rods in a 20 um cube, log-normal lengths, 8 agglomerates, vMF(+z, 5)."""
    base = dict(box=(20, 20, 20), n_rods=500000, len=1.0, len_law=nf.NF_LEN_LOGNORMAL,
                len_sigma=0.3, len_min=0.3, len_max=3.0, radius=0.005, seed=2,
                den_law=nf.NF_DEN_AGGLOM, n_agg=8, agg_sigma=1.5, agg_peak=3.0,
                ort_law=nf.NF_ORT_VMF, kappa=5.0, axis=(0, 0, 1), contact_gap=0.002,
                precision=precision)
    base.update(kw)
    return nf.normalize(nf.default_params(**base))


def tiny_box(precision=nf.NF_PREC_FP64, **kw):
    """Box smaller than a rod: exercises multiple periodic images and self images."""
    base = dict(box=(1.5, 1.5, 1.5), n_rods=60, len=2.0, radius=0.01, seed=11, contact_gap=0.01,
                precision=precision)
    base.update(kw)
    return nf.normalize(nf.default_params(**base))
