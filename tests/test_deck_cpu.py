"""The input deck reader against the reference's own decks: the sod test
case's input (tests/golden/sod.inp = testcases/sod/sod.inp) and the full
configuration the reference wrote back for it with every key and its value
(tests/golden/sod_output.cfg = testcases/sod/output.cfg)."""
from __future__ import annotations

import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def test_reads_the_sod_deck():
    from phd_amd.deck import Deck
    d = Deck.read(os.path.join(GOLD, "sod.inp"))
    assert d.get("GENERAL%VIS") is True and d.get("GENERAL%INIT") == "SOD"
    assert d.get("GENERAL%DIM") == 2 and d.get("GENERAL%NRV") == 2
    assert d.get("TIMESTEPPING%CFL") == 0.5 and d.get("TIMESTEPPING%VIS_DT") is False
    assert d.get("WENO%WEIGHTS") is True and d.get("WENO%DUCROS_TREDUC") == 0.1
    assert d.get("WENO%PERCENT_WEIGHT") == 0.15 and d.get("WENO%ORDER_WENO") == 7
    assert d.get("WENO%SHOCK_SENSOR_TYPE") == "ADAMSSHARIFF"
    assert d.get("MONITORING%FREQ_RESTART") == 500 and d.get("MONITORING%ITMAX") == 2
    assert d.get("MONITORING%TIME_OFFSET") == 7.0968962141250196e-05
    assert d.get("GRID%BCXMIN") == "EXTRAPOLATION" and d.get("GRID%NUMPROCX") == 2
    # keys the file does not set keep read_inp's defaults
    assert d.get("TIMESTEPPING%CPR") == 20.0 and d.get("GENERAL%FORCESUM") is True
    assert d.get("GRID%BCYMIN") == "PERIODIC"


def test_the_written_back_configuration_agrees():
    """Every key of the written-back file parses with its type, and the
    values the input deck set are the ones written back."""
    from phd_amd.deck import KEYS, Deck
    full = Deck.read(os.path.join(GOLD, "sod_output.cfg"))
    inp = Deck.read(os.path.join(GOLD, "sod.inp"))
    assert set(full.raw) == set(KEYS)  # the reference writes every key it knows
    for name in inp.raw:
        a, b = inp.get(name), full.get(name)
        if isinstance(a, float):
            assert b == pytest.approx(a, rel=1e-4)  # written with 5 digits
        else:
            assert a == b, name
    assert full.get("MONITORING%PROBE_NAME") == ["PROBE1", "PROBE2"]
    assert full.get("MONITORING%VOLUME") == [0.0, 0.0, 0.0, 1.0, 1.0, 1.0]
    assert full.get("FILENAME") == "shearlayer.ini"


def test_syntax_rules():
    from phd_amd.deck import Deck, DeckError
    text = ("[WENO]\n"
            "  WEIGHTS = .false.   # comment\n"
            "  SHOCK_SENSOR_TYPE = 'DUC#ROS' ; comment\n"
            "\tPERCENT_WEIGHT = 1.5d-1\n"
            "[GRID]\n"
            "  BCXMIN = 'PERI'\n"
            "  BCXMAX += ''\n")
    d = Deck.parse(text)
    assert d.get("WENO%WEIGHTS") is False
    assert d.get("WENO%SHOCK_SENSOR_TYPE") == "DUC#ROS"
    assert d.get("WENO%PERCENT_WEIGHT") == 0.15
    # an unindented key does not belong to the category
    with pytest.raises(DeckError, match="unknown variable WEIGHTS"):
        Deck.parse("[WENO]\nWEIGHTS = T\n")
    with pytest.raises(DeckError, match="unknown variable"):
        Deck.parse("[WENO]\n  NOT_A_KEY = 1\n")
    with pytest.raises(DeckError, match="category"):
        Deck.parse("WENO\n")
    with pytest.raises(DeckError, match="wrong size"):
        Deck.parse("[MONITORING]\n  VOLUME = 1 2 3\n")
    with pytest.raises(DeckError):
        Deck.parse("[GENERAL]\n  DIM = two\n")


def test_options_refuse_what_is_not_built():
    from phd_amd.deck import Deck, DeckError, options
    sod = Deck.read(os.path.join(GOLD, "sod.inp"))
    with pytest.raises(DeckError) as e:
        options(sod)
    for what in ("ORDER_WENO = 7", "NRV = 2", "NUMPROCX = 2"):
        assert what in str(e.value)
    assert "BCXMIN" not in str(e.value)      # extrapolated faces are built
    from phd_amd.deck import face_conditions
    assert face_conditions(sod) == (1, 1, 0, 0, 0, 0)
    with pytest.raises(DeckError, match="BCYMAX"):
        options(Deck.parse("[GRID]\n  BCYMAX = 'NOREFLECTION'\n"))
    ok = Deck.parse("[GENERAL]\n  DIM = 3\n  NRV = 9\n[WENO]\n  ORDER_WENO = 5\n"
                    "  PERCENT_WEIGHT = -1\n  SHOCK_SENSOR_TYPE = 'DUCROS'\n"
                    "  DUCROS_TREDUC = 0.5\n[REACTING]\n  REACTION = T\n"
                    "[TIMESTEPPING]\n  CFL = 0.4\n  CPR = 10.0\n")
    o = options(ok)
    assert (o.reaction, o.CFL, o.CPR, o.max_rel_weight, o.sensor_type, o.ducros_treduc) == \
        (1, 0.4, 10.0, -1.0, 1, 0.5)
