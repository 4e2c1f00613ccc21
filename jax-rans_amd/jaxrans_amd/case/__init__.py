from .loader import Case, CaseError, GridSpec, apply_override, load_case, parse_case  # noqa: F401
