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
    from phd_amd.deck import (Deck, DeckError, decomposition, face_conditions, library_key,
                              options)
    # the reference's Sod deck runs on the 2-species WENO7 flow-only library,
    # reproducing its 2-rank decomposition along x
    sod = Deck.read(os.path.join(GOLD, "sod.inp"))
    o = options(sod)
    assert (o.vis, o.vis_dt, o.reaction, o.CFL, o.max_rel_weight) == (1, 0, 0, 0.5, 0.15)
    assert library_key(sod) == (2, 4) and decomposition(sod) == (2, 1, 1)
    assert face_conditions(sod) == (1, 1, 0, 0, 0, 0)
    # WENO5 with 2 species has no library; chemistry needs the full one
    with pytest.raises(DeckError, match="NRV = 2 with ORDER_WENO = 5"):
        options(Deck.parse("[GENERAL]\n  NRV = 2\n[WENO]\n  ORDER_WENO = 5\n"))
    with pytest.raises(DeckError, match="REACTION = T with NRV = 9, ORDER_WENO = 7"):
        options(Deck.parse("[GENERAL]\n  NRV = 9\n[WENO]\n  ORDER_WENO = 7\n"
                           "[REACTING]\n  REACTION = T\n"))
    with pytest.raises(DeckError, match="ORDER_WENO = 3"):
        options(Deck.parse("[GENERAL]\n  NRV = 9\n[WENO]\n  ORDER_WENO = 3\n"))
    with pytest.raises(DeckError, match="NUMPROCZ = 2"):
        options(Deck.parse("[GENERAL]\n  DIM = 2\n  NRV = 9\n[WENO]\n  ORDER_WENO = 5\n"
                           "[GRID]\n  NUMPROCZ = 2\n"))
    with pytest.raises(DeckError, match="BCYMAX"):
        options(Deck.parse("[GRID]\n  BCYMAX = 'NOREFLECTION'\n"))
    ok = Deck.parse("[GENERAL]\n  DIM = 3\n  NRV = 9\n[WENO]\n  ORDER_WENO = 5\n"
                    "  PERCENT_WEIGHT = -1\n  SHOCK_SENSOR_TYPE = 'DUCROS'\n"
                    "  DUCROS_TREDUC = 0.5\n[REACTING]\n  REACTION = T\n"
                    "[TIMESTEPPING]\n  CFL = 0.4\n  CPR = 10.0\n")
    o = options(ok)
    assert (o.reaction, o.CFL, o.CPR, o.max_rel_weight, o.sensor_type, o.ducros_treduc) == \
        (1, 0.4, 10.0, -1.0, 1, 0.5)
