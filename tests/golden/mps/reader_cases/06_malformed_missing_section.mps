NAME          MALFORMED
* Missing ROWS and COLUMNS sections completely
RHS
    RHS1      ROW1      10.0
ENDATA