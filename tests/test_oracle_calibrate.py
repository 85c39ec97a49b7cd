"""Calibration's rules and per-row factors (core/calibrate.nim): the oracle's
restatement and the library's host resolution (pseg_calib_resolve, no GPU
work), each held to the values of the reference's tests/tcalibrate.nim, then
to each other on random rules and rows."""
import math

import numpy as np
import pytest

import oracle as O
import processseg_amd as P
from processseg_amd import calibrate as K

NaN = math.nan
D1 = K.descale(1, K.duMicroVPerCount)   # 1e-6 V/count
D2 = K.descale(2, K.duMicroVPerCount)
GEO = K.sensitivity(20, K.suVPerMPerS)
HYD = K.sensitivity(20, K.suVPerBar)    # 2e-4 V/Pa


def close(a, b, rel=1e-12):
    return abs(a - b) <= rel * max(abs(a), abs(b))


def station_row(sensor=2, **kw):
    """stationKeys (tests/tcalibrate.nim:28-34): live, sensor type and
    component `sensor`, channel 3, station 10/201.1."""
    row = dict(status=0, failed=False, sensorType=sensor, component=sensor, channelSet=-1,
               channel=3.0, rcvLine=10.0, rcvPoint=201.0, rcvIndex=1, fsDescale=NaN,
               fsSensitivity=NaN, fsUnit=0, fsPhysical=False)
    row.update(kw)
    return row


def meta_of(rows, scales=False, cols=16):
    n = len(rows)
    fs = None
    if scales:
        fs = [K.FileScale(r["fsDescale"], r["fsSensitivity"], r["fsUnit"], r["fsPhysical"])
              for r in rows]
    failed = [i for i, r in enumerate(rows) if r["failed"]]
    return P.ChunkMeta(n, 2.0, numSamples=[r.get("count", cols) for r in rows],
                       status=[r["status"] for r in rows], failed=failed,
                       sensorType=[r["sensorType"] for r in rows],
                       component=[r["component"] for r in rows],
                       channel=[r["channel"] for r in rows], rcvLine=[r["rcvLine"] for r in rows],
                       rcvPoint=[r["rcvPoint"] for r in rows],
                       rcvIndex=[r["rcvIndex"] for r in rows],
                       channelSet=[r["channelSet"] for r in rows], scales=fs)


def oracle_rule(r: K.CalibrationRule):
    m = r.match
    return dict(sensorType=m.sensorType, component=m.component, channelSet=m.channelSet,
                channel=m.channel,
                station=None if m.station is None else (m.station.line, m.station.point,
                                                        m.station.index),
                descale=r.descale.voltsPerCount, sensitivity=r.sensitivity.voltsPerUnit,
                unit=r.sensitivity.unit, gainDb=r.gainDb)


def both(cal: K.Calibration, rows, scales=False):
    """The oracle's and the library's resolution of `rows` under `cal`:
    two lists of dicts with the same keys."""
    p = O.calib_plan([oracle_rule(r) for r in cal.rules], cal.target, cal.fileFactors,
                     cal.scaleAuxiliary)
    want = [O.calib_resolve_row(p, r, scales) for r in rows]
    got = [dataclass_dict(x) for x in K.initCalibrationPlan(cal).resolve(meta_of(rows, scales))]
    return want, got


def dataclass_dict(x: K.RowCalibration):
    return dict(factor=x.factor, unit=x.unit, status=x.status, descaleFrom=x.descaleFrom,
                sensitivityFrom=x.sensitivityFrom, descaleRule=x.descaleRule,
                sensitivityRule=x.sensitivityRule, gainRule=x.gainRule, conflict=x.conflict,
                mismatch=x.mismatch)


def same(a, b):
    for k in a:
        x, y = a[k], b[k]
        if isinstance(x, float) and math.isnan(x):
            assert isinstance(y, float) and math.isnan(y), (k, a, b)
        else:
            assert x == y, (k, a, b)


def check_each(cal, rows, scales, fn):
    """fn(row result) holds for the oracle and the library alike, and they
    agree field by field."""
    want, got = both(cal, rows, scales)
    for w, g in zip(want, got):
        same(w, g)
    fn(want)
    fn(got)


MATCHES = [K.RuleMatch(station=K.stationKey(10, 201, 0), channel=3),
           K.RuleMatch(station=K.stationKey(10, 201, 1)), K.RuleMatch(channel=3),
           K.RuleMatch(component=2), K.RuleMatch(sensorType=2), K.RuleMatch()]
SENS = [11.0, 12.0, 13.0, 14.0, 15.0, 16.0]


@pytest.mark.parametrize("first", range(len(MATCHES)))
def test_precedence_ladder(first):
    """tcalibrate.nim:36-61: each rank beats the next; no descale anywhere,
    so the row is unresolved but the sensitivity is still chosen."""
    rules = [K.CalibrationRule(match=MATCHES[i], sensitivity=K.sensitivity(SENS[i], K.suVPerMPerS))
             for i in range(first, len(MATCHES))]
    rules.append(K.CalibrationRule(match=K.RuleMatch(sensorType=99), descale=D1))

    def fn(res):
        r = res[0]
        assert r["status"] == K.csUnresolved and r["sensitivityRule"] == 0
        assert r["sensitivityFrom"] == K.fsRule
    check_each(K.initCalibration(rules), [station_row()], False, fn)


def test_precedence_order_of_the_list_and_independent_factors():
    """tcalibrate.nim:62-81: the list's order does not matter; the descale,
    sensitivity and gain resolve independently."""
    rev = [K.CalibrationRule(match=MATCHES[i], sensitivity=K.sensitivity(SENS[i], K.suVPerMPerS))
           for i in range(len(MATCHES) - 1, -1, -1)]
    rev.append(K.CalibrationRule(match=K.RuleMatch(sensorType=99), descale=D1))
    check_each(K.initCalibration(rev), [station_row()], False,
               lambda res: res[0]["sensitivityRule"] == len(MATCHES) - 1 or pytest.fail())
    cal = K.initCalibration([
        K.CalibrationRule(match=K.RuleMatch(station=K.stationKey(10, 201, 0)), sensitivity=GEO),
        K.CalibrationRule(match=K.RuleMatch(sensorType=2), gainDb=20),
        K.CalibrationRule(descale=D1)])

    def fn(res):
        r = res[0]
        assert (r["descaleRule"], r["sensitivityRule"], r["gainRule"]) == (2, 0, 1)
        assert close(r["factor"], 1e-6 / (10 * 20)) and r["unit"] == K.auVelocity
    check_each(cal, [station_row()], False, fn)
    s = K.initCalibrationPlan(cal).resolveInto(meta_of([station_row()]))
    assert s.converted == 1 and s.units[K.auVelocity] == 1


def test_station_index():
    """tcalibrate.nim:82-95: a trace without a receiver index matches a
    station rule for index 1, not one for index 2."""
    row = station_row(rcvIndex=0)
    exact = K.initCalibration([K.CalibrationRule(
        match=K.RuleMatch(station=K.stationKey(10, 201, 1)), descale=D2, sensitivity=GEO)])
    check_each(exact, [row], False, lambda res: (res[0]["status"] == K.csConverted and
                                                close(res[0]["factor"], 2e-6 / 20)) or pytest.fail())
    other = K.initCalibration([K.CalibrationRule(
        match=K.RuleMatch(station=K.stationKey(10, 201, 2)), descale=D2, sensitivity=GEO)])
    check_each(other, [row], False,
               lambda res: res[0]["status"] == K.csUnresolved or pytest.fail())


def test_channel_set_rank():
    """tcalibrate.nim:97-118 on rows shaped like its record: a channel-set
    rule beats a sensor-type rule and loses to a channel rule."""
    rows = [station_row(0, status=3, channelSet=0, channel=1.0)]
    for ch in (2.0, 3.0, 4.0):
        rows.append(station_row(2, channelSet=1, channel=ch))
    for ch in (5.0, 6.0):
        rows.append(station_row(1, channelSet=2, channel=ch))
    cal = K.initCalibration([
        K.CalibrationRule(match=K.RuleMatch(sensorType=2), descale=D1, sensitivity=GEO),
        K.CalibrationRule(match=K.RuleMatch(channelSet=1), descale=D2),
        K.CalibrationRule(match=K.RuleMatch(channel=3),
                          sensitivity=K.sensitivity(40, K.suVPerMPerS)),
        K.CalibrationRule(match=K.RuleMatch(sensorType=1), descale=D1, sensitivity=HYD)])

    def fn(res):
        assert res[0]["status"] == K.csKept
        for r in (1, 2, 3):
            assert res[r]["descaleRule"] == 1 and res[r]["unit"] == K.auVelocity
        assert close(res[1]["factor"], 2e-6 / 20) and close(res[2]["factor"], 2e-6 / 40)
        for r in (4, 5):
            assert res[r]["unit"] == K.auPressure and close(res[r]["factor"], 1e-6 / 2e-4)
    check_each(cal, rows, False, fn)
    s = K.initCalibrationPlan(cal).resolveInto(meta_of(rows))
    assert s.converted == 5 and s.kept == 1 and s.units[K.auPressure] == 2


def file_rows(unit=K.auVelocity):
    return [station_row(fsDescale=3e-6, fsSensitivity=25.0, fsUnit=unit)] * 2


def test_file_factors():
    """tcalibrate.nim:120-172: fallback, first and ignore; agreement within
    0.1 %; the file alone; a unit conflict; a meta without the column."""
    rules = [K.CalibrationRule(match=K.RuleMatch(sensorType=2), descale=D1),
             K.CalibrationRule(sensitivity=GEO)]

    def fallback(res):
        r = res[0]
        assert r["descaleFrom"] == K.fsRule and r["sensitivityFrom"] == K.fsFile and r["conflict"]
        assert close(r["factor"], 1e-6 / 25)
    check_each(K.initCalibration(rules), file_rows(), True, fallback)

    def first(res):
        r = res[0]
        assert r["descaleFrom"] == K.fsFile and r["sensitivityFrom"] == K.fsFile and r["conflict"]
        assert close(r["factor"], 3e-6 / 25)
    check_each(K.initCalibration(rules, fileFactors=K.ffFirst), file_rows(), True, first)

    def ignore(res):
        r = res[0]
        assert r["descaleFrom"] == K.fsRule and r["sensitivityFrom"] == K.fsRule
        assert not r["conflict"] and close(r["factor"], 1e-6 / 20)
    check_each(K.initCalibration(rules, fileFactors=K.ffIgnore), file_rows(), True, ignore)
    agree = K.initCalibration([K.CalibrationRule(
        descale=K.descale(3.002, K.duMicroVPerCount), sensitivity=K.sensitivity(25, K.suVPerMPerS))])
    check_each(agree, file_rows(), True,
               lambda res: (not res[0]["conflict"] and res[0]["descaleFrom"] == K.fsFile)
               or pytest.fail())
    off = K.initCalibration([K.CalibrationRule(descale=K.descale(3.01, K.duMicroVPerCount))])
    check_each(off, file_rows(), True, lambda res: res[0]["conflict"] or pytest.fail())
    check_each(K.initCalibration(), file_rows(), True,
               lambda res: (res[0]["status"] == K.csConverted and
                            close(res[0]["factor"], 3e-6 / 25) and
                            res[0]["unit"] == K.auVelocity) or pytest.fail())
    check_each(K.initCalibration([K.CalibrationRule(sensitivity=GEO)]), file_rows(K.auPressure),
               True, lambda res: (res[0]["conflict"] and res[0]["unit"] == K.auPressure)
               or pytest.fail())
    check_each(K.initCalibration(), file_rows(), False,
               lambda res: res[0]["status"] == K.csUnresolved or pytest.fail())


def test_already_physical():
    """tcalibrate.nim:174-191."""
    row = [station_row(1, fsDescale=0.125, fsUnit=K.auPressure, fsPhysical=True)]
    check_each(K.initCalibration(), row, True,
               lambda res: (res[0]["factor"] == 0.125 and res[0]["unit"] == K.auPressure and
                            not res[0]["conflict"]) or pytest.fail())
    with_rule = [K.CalibrationRule(descale=D1, sensitivity=HYD)]
    check_each(K.initCalibration(with_rule), row, True,
               lambda res: (res[0]["factor"] == 0.125 and res[0]["conflict"]) or pytest.fail())
    check_each(K.initCalibration(with_rule, fileFactors=K.ffIgnore), row, True,
               lambda res: (close(res[0]["factor"], 1e-6 / 2e-4) and not res[0]["conflict"])
               or pytest.fail())
    check_each(K.initCalibration(target=K.ctVolts), row, True,
               lambda res: res[0]["status"] == K.csUnresolved or pytest.fail())


def status_rows():
    """tcalibrate.nim:193-211: live, live accelerometer, dead with 10 live
    samples, dead accelerometer, auxiliary, failed."""
    rows = [station_row(2) for _ in range(6)]
    rows[1]["sensorType"] = 7
    rows[2]["status"] = 2
    rows[2]["count"] = 10
    rows[3]["status"] = 2
    rows[3]["sensorType"] = 7
    rows[4]["status"] = 3
    rows[4]["sensorType"] = 0
    rows[5]["failed"] = True
    return rows


def test_rows_by_status():
    """tcalibrate.nim:193-245: statuses, counts, auxiliary rows in volts
    (never a sensitivity), scaleAuxiliary off, the volts target."""
    rows = status_rows()
    cal = K.initCalibration([K.CalibrationRule(match=K.RuleMatch(sensorType=2), sensitivity=GEO),
                             K.CalibrationRule(descale=D1)])

    def fn(res):
        st = [r["status"] for r in res]
        assert st == [K.csConverted, K.csUnresolved, K.csConverted, K.csKept, K.csConverted,
                      K.csKept]
        assert res[4]["unit"] == K.auVolt and res[4]["sensitivityFrom"] == K.fsNone
        assert res[5]["factor"] == 1 and math.isnan(res[1]["factor"])
        assert close(res[0]["factor"], 1e-6 / 20)
    check_each(cal, rows, False, fn)
    s = K.initCalibrationPlan(cal).resolveInto(meta_of(rows))
    assert (s.converted, s.unresolved, s.kept, s.noSensitivity) == (3, 1, 2, 1)
    assert s.units[K.auVelocity] == 2 and s.units[K.auVolt] == 1 and s.units[K.auRecorded] == 2
    one = K.initCalibration([K.CalibrationRule(descale=D1, sensitivity=GEO)])
    check_each(one, rows, False, lambda res: (res[4]["unit"] == K.auVolt and
                                             res[4]["sensitivityRule"] == -1 and
                                             res[4]["factor"] == 1e-6) or pytest.fail())
    no_aux = K.initCalibration([K.CalibrationRule(descale=D1, sensitivity=GEO)],
                               scaleAuxiliary=False)
    check_each(no_aux, rows, False, lambda res: res[4]["status"] == K.csKept or pytest.fail())
    volts = K.initCalibration([K.CalibrationRule(descale=D1)], target=K.ctVolts)
    check_each(volts, rows, False, lambda res: (res[1]["factor"] == 1e-6 and
                                               res[0]["sensitivityFrom"] == K.fsNone)
               or pytest.fail())
    vs = K.initCalibrationPlan(volts).resolveInto(meta_of(rows))
    assert vs.converted == 5 and vs.units[K.auVolt] == 5


def test_gain_and_mismatch():
    """tcalibrate.nim:247-285."""
    hyd_row = [station_row(1)]
    check_each(K.initCalibration([K.CalibrationRule(descale=D1, sensitivity=HYD, gainDb=12)]),
               hyd_row, False,
               lambda res: (close(res[0]["factor"], 1e-6 / (10 ** 0.6 * 2e-4)) and
                            not res[0]["mismatch"]) or pytest.fail())
    check_each(K.initCalibration([K.CalibrationRule(descale=D1, sensitivity=GEO)]), hyd_row,
               False, lambda res: (res[0]["mismatch"] and res[0]["unit"] == K.auVelocity)
               or pytest.fail())
    plain = [dict(station_row(0), sensorType=0, component=0, channel=NaN, rcvLine=NaN,
                  rcvPoint=NaN, rcvIndex=0, fsDescale=1e-6, fsUnit=K.auPressure)]
    check_each(K.initCalibration([K.CalibrationRule(sensitivity=GEO)]), plain, True,
               lambda res: (res[0]["mismatch"] and res[0]["status"] == K.csConverted)
               or pytest.fail())
    check_each(K.initCalibration([K.CalibrationRule(sensitivity=HYD)]), plain, True,
               lambda res: (not res[0]["mismatch"]) or pytest.fail())
    acc = [station_row(7)]
    check_each(K.initCalibration([K.CalibrationRule(descale=D1,
                                                    sensitivity=K.sensitivity(1, K.suVPerG))]),
               acc, False, lambda res: (not res[0]["mismatch"] and
                                        res[0]["unit"] == K.auAcceleration) or pytest.fail())
    check_each(K.initCalibration([K.CalibrationRule(descale=D1, sensitivity=GEO)]), acc, False,
               lambda res: res[0]["mismatch"] or pytest.fail())
    check_each(K.initCalibration([K.CalibrationRule(descale=D1, sensitivity=HYD, gainDb=-20)]),
               hyd_row, False, lambda res: close(res[0]["factor"], 1e-6 * 10 / 2e-4) or pytest.fail())


def test_units():
    """core/units.nim's examples: 0.7 V/(in/s), 20 V/bar = 2e-4 V/Pa,
    0.5 µV/count, and the open-circuit divider 20 · 1600 / 2000."""
    assert abs(K.sensitivity(0.7, K.suVPerInPerS).voltsPerUnit - 0.7 / 0.0254) < 1e-12
    assert K.sensitivity(20, K.suVPerBar).voltsPerUnit == 2e-4
    assert K.descale(0.5, K.duMicroVPerCount).voltsPerCount == 5e-7
    s = K.openCircuitSensitivity(20, K.suVPerMPerS, coilOhms=400, shuntOhms=1600)
    assert abs(s.voltsPerUnit - 16) < 1e-12 and s.form == K.snOpenCircuit
    assert close(K.sensitivity(-192, K.suDbReVPerMicroPa).voltsPerUnit, 10 ** (-9.6) * 1e6)
    with pytest.raises(ValueError):
        K.sensitivity(0, K.suVPerMPerS)
    with pytest.raises(ValueError):
        K.descale(-1, K.duVPerCount)


@pytest.mark.parametrize("rules,msg", [
    ([K.CalibrationRule()], "rule 0: gives no factor"),
    ([K.CalibrationRule(descale=D1), K.CalibrationRule(sensitivity=GEO)], "rule 1: has no criteria"),
    ([K.CalibrationRule(K.RuleMatch(sensorType=2), descale=D1),
      K.CalibrationRule(K.RuleMatch(sensorType=2), sensitivity=GEO)],
     "rule 1: matches the same rows as rule 0"),
    ([K.CalibrationRule(gainDb=math.inf)], "rule 0: gainDb must be finite"),
    ([K.CalibrationRule(gainDb=1e6)], "rule 0: gainDb = "),
    ([K.CalibrationRule(K.RuleMatch(sensorType=256), descale=D1)], "rule 0: sensorType must be"),
    ([K.CalibrationRule(K.RuleMatch(component=-2), descale=D1)], "rule 0: component must be"),
    ([K.CalibrationRule(K.RuleMatch(channelSet=-3), descale=D1)], "rule 0: channelSet must be"),
    ([K.CalibrationRule(K.RuleMatch(channel=math.inf), descale=D1)], "rule 0: channel must be"),
    ([K.CalibrationRule(K.RuleMatch(station=K.stationKey(math.nan, 1)), descale=D1)],
     "rule 0: the station's line and point"),
])
def test_plan_refusals(rules, msg):
    """initCalibrationPlan's ValueErrors (core/calibrate.nim:328-352), in the
    library and the oracle."""
    with pytest.raises(ValueError) as e:
        K.initCalibrationPlan(K.initCalibration(rules))
    assert str(e.value).startswith(msg), str(e.value)
    with pytest.raises(ValueError) as e2:
        O.calib_plan([oracle_rule(r) for r in rules])
    assert str(e2.value).startswith(msg.split(":")[0]), str(e2.value)


def test_channel_and_station_rounding():
    """Criteria compare to the hundredth (core/calibrate.nim:225, 256-264):
    channel 3.004 matches a channel-3 rule, 3.006 does not."""
    cal = K.initCalibration([K.CalibrationRule(K.RuleMatch(channel=3), descale=D1,
                                               sensitivity=GEO)])
    check_each(cal, [station_row(channel=3.004), station_row(channel=3.006)], False,
               lambda res: (res[0]["status"] == K.csConverted and
                            res[1]["status"] == K.csUnresolved) or pytest.fail())


def test_random_rules_and_rows():
    """The library's resolution equals the oracle's on 40 random rule sets
    over 200 random rows each (criteria drawn from a small domain so rules
    collide and rank against each other; file scales on half the sets)."""
    rng = np.random.default_rng(7)
    units = [K.suVPerMPerS, K.suVPerBar, K.suVPerG]
    for case in range(40):
        rules, keys = [], set()
        while len(rules) < rng.integers(1, 8):
            m = K.RuleMatch()
            if rng.random() < 0.4:
                m.sensorType = int(rng.integers(0, 4))
            if rng.random() < 0.3:
                m.component = int(rng.integers(0, 3))
            if rng.random() < 0.2:
                m.channelSet = int(rng.integers(0, 2))
            if rng.random() < 0.2:
                m.channel = float(rng.integers(1, 4))
            if rng.random() < 0.2:
                m.station = K.stationKey(10, float(rng.integers(200, 202)), int(rng.integers(0, 3)))
            k = (m.sensorType, m.component, m.channelSet, m.channel if m.channel == m.channel else -9,
                 m.station)
            if k in keys:
                continue
            keys.add(k)
            r = K.CalibrationRule(m)
            if rng.random() < 0.6:
                r.descale = K.descale(float(rng.uniform(0.1, 5)), K.duMicroVPerCount)
            if rng.random() < 0.6:
                r.sensitivity = K.sensitivity(float(rng.uniform(1, 50)),
                                              units[int(rng.integers(0, 3))])
            if rng.random() < 0.3:
                r.gainDb = float(rng.uniform(-20, 40))
            if not (r.descale.isSet() or r.sensitivity.isSet() or r.gainDb == r.gainDb):
                r.gainDb = 6.0
            rules.append(r)
        cal = K.initCalibration(rules, target=int(rng.integers(0, 2)),
                                fileFactors=int(rng.integers(0, 3)),
                                scaleAuxiliary=bool(rng.integers(0, 2)))
        scales = case % 2 == 0
        rows = []
        for _ in range(200):
            rows.append(dict(
                status=int(rng.integers(0, 4)), failed=bool(rng.random() < 0.05),
                sensorType=int(rng.integers(0, 9)), component=int(rng.integers(0, 3)),
                channelSet=int(rng.integers(-1, 2)),
                channel=float(rng.integers(1, 4)) if rng.random() < 0.8 else NaN,
                rcvLine=10.0 if rng.random() < 0.9 else NaN,
                rcvPoint=float(rng.integers(200, 202)), rcvIndex=int(rng.integers(0, 3)),
                fsDescale=float(rng.choice([NaN, 1e-6, 3e-6])),
                fsSensitivity=float(rng.choice([NaN, 20.0, 25.0])),
                fsUnit=int(rng.choice([0, 2, 4])), fsPhysical=bool(rng.random() < 0.1)))
        try:
            want, got = both(cal, rows, scales)
        except ValueError:
            continue
        for w, g in zip(want, got):
            same(w, g)
