"""`python -m ttc --perm=... --size=...`: the ttc command line."""
from ttc_amd.ttc import main

if __name__ == "__main__":
    main()
