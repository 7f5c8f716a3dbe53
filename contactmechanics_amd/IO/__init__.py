from .NetCDF import NetCDFContainer, NetCDFContainerFrame, open  # noqa: F401,A004
