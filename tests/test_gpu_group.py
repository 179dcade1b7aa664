"""Ragged groups (gbkhip_group_*) and the autotuner: galaxies of different
sizes and models (thin smdisk on the direct pipeline, thick smdisk and
mcdisk on the direct or padded-cube one, as timed) evaluated in one
launch.  Each member's chi^2 in the group equals its own evaluation (graph
and direct launches) bit for bit; smdisk members' cubes match the FP64
oracle within 1e-5 + amb."""
import numpy as np
import pytest

import hiputil as H
import pipeline_cases as P

pytestmark = pytest.mark.gpu


def _same(p, a, b):
    """Bit-identical (every sum of every pipeline has a fixed order)."""
    return np.array_equal(a, b)


def _galaxies(drv):
    thin_a = dict(P.c1_params(), size=(48, 40, 32))
    thin_b = dict(P.c1_params(), size=(30, 36, 24), scale=(3, 3, 1),
                  psf_fwhm=2.0)
    thick = dict(P.c1_params(), thin=False, size=(32, 32, 24), size_z=24)
    mc = dict(P.c1_params(), size=(32, 32, 24), size_z=24, cflux=0.01)
    out = []
    for kind, p in (('sm', thin_a), ('sm', thin_b), ('sm', thick),
                    ('mc', mc)):
        if kind == 'mc':
            model, base = P.build_mc_product(drv, p, cflux=p['cflux'])
        else:
            model, base = P.build_product(drv, p)
        ref = model.model_h(base)['scube']['d']
        rng = np.random.default_rng(len(out))
        mock = (ref + rng.normal(0, 0.01, ref.shape)).astype(np.float32)
        sets = [dict(base, incl=i, posa=a) for i, a in
                ((40, 20), (55, 140), (70, 260))]
        out.append((kind, p, model, base, mock, sets))
    return out


@pytest.fixture(scope='module')
def group():
    from gbkfit_amd.batch import RaggedGroup, autotune
    drv = H.driver()
    gals = _galaxies(drv)
    evs, reports = [], []
    for kind, p, model, base, mock, sets in gals:
        ev, rep = autotune(drv, model.dmodel(), model.gmodel(), mock,
                           sets + [base], error=0.01, max_batches=(2, 4),
                           reps=2, cache=False)
        evs.append(ev)
        reports.append(rep)
    g = RaggedGroup(evs)
    yield drv, gals, evs, reports, g
    g.close()
    for ev in evs:
        ev.close()


def test_autotuner_choices(group):
    _, gals, evs, reports, _ = group
    for (kind, p, *_), ev, rep in zip(gals, evs, reports):
        path, mb, tiles = rep['choice']
        assert ev.info()['pipeline'] == path
        assert rep['trials'] and min(t['s_per_set'] for t in
                                     rep['trials']) == rep['s_per_set']
        tried = {t['path'] for t in rep['trials']}
        if kind == 'mc' or not p['thin']:
            assert tried == {'direct', 'padded'}, tried
        print(kind, p['size'], p['scale'], rep['choice'],
              '%.1f us/set' % (1e6 * rep['s_per_set']))


@pytest.mark.parametrize('use_graph', [False, True])
def test_group_equals_members_alone(group, use_graph):
    _, gals, evs, _, g = group
    chi2 = g.evaluate([sets for *_, sets in gals], use_graph=use_graph)
    for (kind, p, *_, sets), ev, c in zip(gals, evs, chi2):
        alone = ev.evaluate(sets, use_graph=False)
        assert _same(p, alone, c), (kind, p['size'], alone, c)


def test_group_ragged_set_counts_and_replay(group):
    """Members with different set counts (one empty); graph replays after
    eager work give the same results."""
    drv, gals, evs, _, g = group
    lists = [gals[0][5], gals[1][5][:1], [], gals[3][5][:2]]
    first = g.evaluate(lists, use_graph=True)
    x = drv.mem_alloc_d((1 << 20,), np.float32)
    drv.mem_fill(x, 1.0)
    again = g.evaluate(lists, use_graph=True)
    for a, b, l, ev, gal in zip(first, again, lists, evs, gals):
        assert len(a) == len(l)
        assert _same(gal[1], a, b)
        if l:
            assert _same(gal[1], a, ev.evaluate(l, use_graph=False))


def test_group_cubes_match_oracle(group):
    _, gals, evs, _, g = group
    chi2, cubes = g.evaluate([sets for *_, sets in gals], want_cubes=True,
                             use_graph=False)
    for (kind, p, model, base, mock, sets), cb in zip(gals, cubes):
        for i, s in enumerate(sets):
            if kind == 'mc':
                plain = model.model_h(s)['scube']['d']
                assert H.normwise(cb[i], plain) < 1e-5
                continue
            q = dict(p, incl=s['incl'], posa=s['posa'])
            ref, amb, _ = P.oracle_cube(q)
            assert P.excess_error(cb[i], ref, amb) < 1e-5, (p['size'], i)
