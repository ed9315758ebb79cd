"""`python -m ttc_amd --perm=... --size=...`: the ttc command line."""
from .ttc import main

if __name__ == "__main__":
    main()
