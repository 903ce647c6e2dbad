"""Global verbosity flag and [TENEX] status printing (tenex/_log.py:1-24)."""
_VERBOSE = False


def set_verbose(flag: bool) -> None:
    global _VERBOSE
    _VERBOSE = bool(flag)


def is_verbose() -> bool:
    return _VERBOSE


def vprint(*args, **kwargs) -> None:
    if _VERBOSE:
        print(*args, **kwargs, flush=True)
