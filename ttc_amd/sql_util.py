"""Database lookups for the tuning cache (counterpart of ttc/sql_util.py).

Same queries and semantics as the reference (size row, permutation row,
measurement row keyed by types, beta == 0 or not, thread count, hotA/hotB
and architecture; best variant by bandwidth), written with bound parameters.
The GPU variant rows carry ttc_amd's kernel configuration in extra
columns (see ttc.createTables).
"""
import json
import sqlite3

OKGREEN = '\033[92m'
FAIL = '\033[91m'
WARNING = '\033[93m'
ENDC = '\033[0m'

MAX_DIM = 10  # columns size0..size9 / idx0..idx9 (ttc/ttc.py:229-262)


def _where_dims(prefix, values):
    parts, args = [], []
    for i in range(MAX_DIM):
        if i < len(values):
            parts.append("%s%d = ?" % (prefix, i))
            args.append(int(values[i]))
        else:
            parts.append("%s%d is NULL" % (prefix, i))
    return " and ".join(parts), args


def getSizeId(cursor, size):
    """ttc/sql_util.py:8-29."""
    w, a = _where_dims("size", size)
    try:
        cursor.execute("select size_id from size where %s;" % w, a)
    except sqlite3.Error as e:
        print(FAIL + "[TTC] ERROR (sql):", e.args[0], ENDC)
        return -1
    r = cursor.fetchall()
    return r[0][0] if r else -1


def getPermId(cursor, perm):
    """ttc/sql_util.py:32-53."""
    w, a = _where_dims("idx", perm)
    try:
        cursor.execute("select permutation_id from permutation where %s;" % w, a)
    except sqlite3.Error as e:
        print(FAIL + "[TTC] ERROR (sql):", e.args[0], ENDC)
        return -1
    r = cursor.fetchall()
    return r[0][0] if r else -1


def getMeasurementId(cursor, sizeId, permId, floatTypeA, floatTypeB, beta, numThreads, hotA, hotB, architecture):
    """ttc/sql_util.py:55-77 (beta only as zero / non-zero)."""
    betaStr = "beta != 0" if beta != 0 else "beta = 0"
    q = ("select measurement_id from measurements where size_id = ? and permutation_id = ? and floatTypeA = ? "
         "and floatTypeB = ? and %s and numThreads >= ? and hotA = ? and hotB = ? and architecture = ? "
         "order by measurement_id desc;" % betaStr)
    try:
        cursor.execute(q, (sizeId, permId, floatTypeA, floatTypeB, numThreads, hotA, hotB, architecture))
    except sqlite3.Error as e:
        print(FAIL + "[TTC] ERROR (sql):", e.args[0], ENDC)
        return -1
    r = cursor.fetchall()
    return r[0][0] if r else -1


def getBestImplementation(cursor, measurementId):
    """(blockA, blockB, loopOrder_id, prefetchDistance, bandwidth) of the
    fastest variant (ttc/sql_util.py:79-95)."""
    try:
        cursor.execute("select blockA, blockB, loopOrder_id, prefetchDistance, bandwidth from variant "
                       "where measurement_id = ? order by bandwidth DESC;", (measurementId,))
    except sqlite3.Error as e:
        print(FAIL + "[TTC] ERROR (sql):", e.args[0], ENDC)
        return -1
    r = cursor.fetchall()
    return tuple(r[0]) if r else -1


def getBestConfig(cursor, measurementId):
    """(config dict, bandwidth) of the fastest GPU variant, or -1."""
    try:
        cursor.execute("select config, bandwidth from variant where measurement_id = ? and config is not NULL "
                       "order by bandwidth DESC;", (measurementId,))
    except sqlite3.Error as e:
        print(FAIL + "[TTC] ERROR (sql):", e.args[0], ENDC)
        return -1
    r = cursor.fetchall()
    if not r:
        return -1
    return json.loads(r[0][0]), r[0][1]


def getLoopPermFrom(cursor, loopId, dim):
    """ttc/sql_util.py:98-121."""
    cols = ", ".join("idx%d" % i for i in range(dim))
    try:
        cursor.execute("select %s from loopOrder where loopOrder_id = ?;" % cols, (loopId,))
    except sqlite3.Error as e:
        print(FAIL + "[TTC] ERROR (sql):", e.args[0], ENDC)
        return -1
    r = cursor.fetchall()
    return list(r[0][:dim]) if r else -1
