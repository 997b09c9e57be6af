"""BackendError {InvalidParameters, InvalidState} (backend/src/errors.rs:16-21).

Native statuses map as the reference Rust glue maps them
(backend/src/ascend910b3/status.rs:23-44): kind 1 -> InvalidParameters,
every other failure -> InvalidState.
"""


class BackendError(Exception):
    def __init__(self, message: str = "", native_code: int = 0):
        super().__init__(message)
        self.native_code = native_code


class InvalidParameters(BackendError):
    pass


class InvalidState(BackendError):
    pass


def check(status, what: str) -> None:
    if status.kind == 0 and status.native_code == 0:
        return
    if status.kind == 1:
        raise InvalidParameters(f"{what}: invalid parameters", status.native_code)
    raise InvalidState(f"{what}: status kind {status.kind} native {status.native_code}", status.native_code)
