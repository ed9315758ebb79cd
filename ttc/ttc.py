"""ttc.ttc: the generator's driver and command line (ttc/ttc.py in the
reference), provided by ttc_amd.ttc."""
from ttc_amd.ttc import *  # noqa: F401,F403
from ttc_amd.ttc import (createTables, database_path, generateTransposition, getTransposeName,  # noqa: F401
                         insertIntoLoopOrder, insertIntoMeasurements, insertIntoPermutation, insertIntoSize,
                         insertIntoVariant, main, parse_args, printEpilog, printHelp)

if __name__ == "__main__":
    main()
